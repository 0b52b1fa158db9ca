// Store helpers for outputs whose element type differs from the FP64 compute type.
#pragma once
#include "common.hpp"
PC_DEV void store_vt(float* p, double v) { *p = float(v); }
PC_DEV void store_vt(double* p, double v) { *p = v; }
PC_DEV void store_vt(cfloat* p, cdouble v) { p->x = float(v.x); p->y = float(v.y); }
PC_DEV void store_vt(cdouble* p, cdouble v) { *p = v; }
