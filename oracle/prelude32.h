/* Pre-included when compiling the reference sources as an FP32 build
 * (oracle/Makefile): the libc headers keep their double prototypes, every
 * other `double` becomes `float`, and -fsingle-precision-constant makes the
 * literals float.  f64 stays a real double for the shim's API. */
#include <math.h>
#include <string.h>
#include <stdlib.h>
#include <stddef.h>
#include <stdio.h>
typedef double f64;
#define double float
