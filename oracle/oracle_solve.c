/* oracle_solve.c -- CPU restatement of the implicit path (test infrastructure only). */
#include "oracle.h"
