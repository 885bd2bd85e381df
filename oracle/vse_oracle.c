/*
 * vse_oracle.c -- FP32 and FP64 CPU oracle of the reference time loop.
 * TEST INFRASTRUCTURE: loaded only by tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg.  See vse_oracle.h for the reference map.
 */
#include <stddef.h>
#include "vse_oracle.h"

#define REAL float
#define SFX(name) name##_f32
#include "vse_oracle_body.h"
#undef REAL
#undef SFX

#define REAL double
#define SFX(name) name##_f64
#include "vse_oracle_body.h"
#undef REAL
#undef SFX

int oracle_abi_version(void) { return 2; }
size_t oracle_params_size(void) { return sizeof(OracleParams); }
