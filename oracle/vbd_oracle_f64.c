/* FP64 instantiation of the CPU oracle (test infrastructure). */
#include "vbd_oracle.h"
#define REAL double
#define SFX f64
#include "vbd_oracle_impl.h"
