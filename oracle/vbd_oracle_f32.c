/* FP32 instantiation of the CPU oracle (test infrastructure). */
#include "vbd_oracle.h"
#define REAL float
#define SFX f32
#define ORACLE_IS_F32 1
#include "vbd_oracle_impl.h"
