// pg_adjcheck.cpp -- host (FP64) evaluation of the per-cell functions of
// pg_device.h and of their adjoints in pg_adjoint_fn.h, for derivative
// checks on the CPU.  Built as a separate library (libpg_adjcheck.so); the
// product library does not contain it.
#include <cstring>

#include "pg_adjoint_fn.h"

using namespace pg;

extern "C" {

// Number of inputs and outputs of function `fn` (0 on an unknown id).
int adjc_shape(int fn, int* nin, int* nout) {
    static const int shape[][2] = {{4, 1}, {4, 1}, {12, 2}, {8, 1}, {15, 5}, {15, 5}, {13, 3}, {13, 3}, {10, 4}, {9, 4}};
    if (fn < 0 || fn >= (int)(sizeof(shape) / sizeof(shape[0]))) return 0;
    *nin = shape[fn][0];
    *nout = shape[fn][1];
    return 1;
}

// y = f(x); lx = (df/dx)^T ly.  fn: 0 aveL, 1 aveR, 2 slopes, 3 flatten5,
// 4 trace_right, 5 trace_left, 6 trace_right_c, 7 trace_left_c, 8 hllc_iso,
// 9 hllc_c.  Inputs that are constants of the sweep (sound speeds, dt, 1/dx,
// weights) get lx = 0.
int adjc_eval(int fn, const double* x, const double* ly, double* y, double* lx) {
    int nin, nout;
    if (!adjc_shape(fn, &nin, &nout)) return -1;
    std::memset(lx, 0, sizeof(double) * nin);
    switch (fn) {
        case 0:
            y[0] = aveL(x[0], x[1], x[2], x[3]);
            aveL_adj(x[0], x[1], x[2], x[3], ly[0], lx[0], lx[1], lx[2], lx[3]);
            break;
        case 1:
            y[0] = aveR(x[0], x[1], x[2], x[3]);
            aveR_adj(x[0], x[1], x[2], x[3], ly[0], lx[0], lx[1], lx[2], lx[3]);
            break;
        case 2: {
            slopes(x[0], x[1], x[2], x[3], x[4], x[5], x[6], x[7], x[8], x[9], x[10], x[11], y[0], y[1]);
            slopes_adj5(x, x[5], x[6], x[7], x[8], x[9], x[10], x[11], ly[0], ly[1], lx, lx[11]);
            break;
        }
        case 3:
            y[0] = flatten5(x[0], x[1], x[2], x[3], x[4], x[5], x[6], x[7], 1.0);
            flatten5_adj(x[0], x[1], x[2], x[3], x[4], x[5], x[6], x[7], ly[0], lx);
            break;
        case 4:
        case 5: {
            const double* in = x + 3;
            if (fn == 4)
                trace_right(x[0], x[1], x[2], in[0], in[1], in[2], in[3], in[4], in[5], in[6], in[7], in[8], in[9],
                            in[10], in[11], y[0], y[1], y[2], y[3], y[4]);
            else
                trace_left(x[0], x[1], x[2], in[0], in[1], in[2], in[3], in[4], in[5], in[6], in[7], in[8], in[9],
                           in[10], in[11], y[0], y[1], y[2], y[3], y[4]);
            if (fn == 4) trace_right_adj(x[0], x[1], x[2], in, ly[0], ly[1], ly[2], ly[3], ly[4], lx + 3);
            else trace_left_adj(x[0], x[1], x[2], in, ly[0], ly[1], ly[2], ly[3], ly[4], lx + 3);
            break;
        }
        case 6:
        case 7: {
            const double* in = x + 4;
            if (fn == 6)
                trace_right_c(x[0], x[1], x[2], x[3], in[0], in[1], in[2], in[3], in[4], in[5], in[6], in[7], in[8],
                              y[0], y[1], y[2]);
            else
                trace_left_c(x[0], x[1], x[2], x[3], in[0], in[1], in[2], in[3], in[4], in[5], in[6], in[7], in[8],
                             y[0], y[1], y[2]);
            if (fn == 6) trace_right_c_adj(x[0], x[1], x[2], x[3], in, ly[0], ly[1], ly[2], lx + 4);
            else trace_left_c_adj(x[0], x[1], x[2], x[3], in, ly[0], ly[1], ly[2], lx + 4);
            break;
        }
        case 8: {
            hllc_iso(x[0], x[1], x[2], x[3], x[4], x[5], x[6], x[7], x[8], x[9], y[0], y[1], y[2], y[3]);
            hllc_iso_adj(x[0], x[1], x + 2, ly[0], ly[1], ly[2], ly[3], lx + 2, lx[0], lx[1]);
            break;
        }
        case 9: {
            hllc_c(x[0], x[1], x[2], x[3], x[4], x[5], x[6], x[7], x[8], y[0], y[1], y[2], y[3]);
            hllc_c_adj(x[0], x[1], x[2], x + 3, ly[0], ly[1], ly[2], ly[3], lx + 3);
            break;
        }
    }
    return 0;
}

}  // extern "C"
