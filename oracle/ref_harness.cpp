// ref_harness.cpp -- extern "C" entry points over the reference's own header-only sources,
// compiled where they lie (src/mesh/kernel.h with KERNEL_FUNCTION set on the command line,
// src/math_types/{vec3,mat3,symmetric_tensor2}.h), into oracle/_ref/.  Test infrastructure:
// the oracle's kernel and 3x3 inverse are compared against these, bit for bit.
#include "src/math_types/vec3.h"
#include "src/math_types/mat3.h"
#include "src/math_types/symmetric_tensor2.h"
#include "src/mesh/kernel.h"

extern "C" {

int ref_kernel_function(void) { return KERNEL_FUNCTION; }

void ref_kernel(double r, double h, double *wk, double *dwk)
{
    double hinv, hinv3, hinv4;
    kernel_hinv(h, &hinv, &hinv3, &hinv4);
    double u = r * hinv;
    *wk = 0;
    *dwk = 0;
    kernel_main(u, hinv3, hinv4, wk, dwk, 0);
}

double ref_mat3_invert(const double *T, double *Tinv)
{
    Mat3<double> m, inv;
    for (int a = 0; a < 3; a++)
        for (int b = 0; b < 3; b++) m[a][b] = T[3 * a + b];
    double d = m.invert(inv);
    for (int a = 0; a < 3; a++)
        for (int b = 0; b < 3; b++) Tinv[3 * a + b] = inv[a][b];
    return d;
}

double ref_mat3_frobenius_norm_sq(const double *T)
{
    Mat3<double> m;
    for (int a = 0; a < 3; a++)
        for (int b = 0; b < 3; b++) m[a][b] = T[3 * a + b];
    return m.frobenius_norm_sq();
}

void ref_sym_matvec(const double *s6, const double *v, double *out)
{
    SymmetricTensor2<double> s;
    for (int k = 0; k < 6; k++) s.data[k] = s6[k];
    Vec3<double> x{v[0], v[1], v[2]};
    Vec3<double> y = s.matvec(x);
    out[0] = y[0]; out[1] = y[1]; out[2] = y[2];
}

}
