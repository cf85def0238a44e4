/* Pre-included when compiling the reference sources in FP64 (oracle/Makefile).
 * Declares f64 for the shim's double API. */
#include <math.h>
#include <string.h>
#include <stdlib.h>
#include <stddef.h>
#include <stdio.h>
typedef double f64;
