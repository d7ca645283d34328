/* Declarations of the parts of R's C API that r/r_profit_amd.cpp uses, for a
 * syntax-only compile of that file where R is not installed: the check is
 * about its calls into include/profit_amd.h.  Names and signatures follow R's
 * documented API (Writing R Extensions, section 5); nothing here is built into
 * a program. */
#pragma once
#include <stddef.h>
#include "Rinternals.h"
