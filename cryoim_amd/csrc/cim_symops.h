/*
 * cim_symops.h -- the 60 rotations of the icosahedral group in the
 * "2-fold axes on x, y, z" setting used by cryoIM
 * (icosprocess/SRC/icos2f_sym_mat.f90:1-248), generated in double
 * precision instead of read from a 14-digit table.
 *
 * The group is the closure of two generators: a 72 degree rotation about
 * the 5-fold axis (1, 0, phi) and a 180 degree rotation about z.  The
 * closure is enumerated breadth first, which fixes a canonical order; the
 * table cim_icos60_order[] then lists, for the reference's operator k, the
 * index of the same rotation in that canonical order.  The reference's
 * operator order matters where the first operator satisfying a test is
 * taken (return_icos_asymmetry_unit.f90:17-28), so it is kept.
 *
 * Plain C, header only: included by the oracle (C) and by the HIP
 * library's host code (C++).
 */
#ifndef CIM_SYMOPS_H
#define CIM_SYMOPS_H

#include <math.h>
#include <string.h>

static const unsigned char cim_icos60_order[60] = {
    0, 17, 24, 8, 10, 16, 49, 30, 7, 34, 57, 47, 13, 36, 58, 48, 14, 12, 46, 33,
    9, 1, 3, 6, 11, 41, 38, 35, 19, 28, 22, 39, 43, 20, 37, 42, 31, 44, 40, 45,
    32, 18, 5, 15, 4, 23, 50, 51, 25, 2, 21, 55, 53, 56, 26, 54, 52, 29, 59, 27};

/* Rotation by ang radians about the unit vector along (ax, ay, az), row-major. */
static void cim_axis_rotation(double ax, double ay, double az, double ang, double* R)
{
    double n = sqrt(ax * ax + ay * ay + az * az);
    double x = ax / n, y = ay / n, z = az / n;
    double K[9] = {0, -z, y, z, 0, -x, -y, x, 0};
    double K2[9];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
            K2[3 * i + j] = K[3 * i] * K[j] + K[3 * i + 1] * K[3 + j] + K[3 * i + 2] * K[6 + j];
    double s = sin(ang), c = 1.0 - cos(ang);
    for (int i = 0; i < 9; ++i) R[i] = (i % 4 == 0 ? 1.0 : 0.0) + s * K[i] + c * K2[i];
}

/* Canonical (breadth-first) enumeration of the 60 group elements. */
static void cim_icos60_canonical(double* out /* 60*9 */)
{
    const double phi = (1.0 + sqrt(5.0)) / 2.0;
    const double pi = 3.14159265358979323846;
    double gen[2][9];
    cim_axis_rotation(1.0, 0.0, phi, 2.0 * pi / 5.0, gen[0]);
    cim_axis_rotation(0.0, 0.0, 1.0, pi, gen[1]);
    int count = 1, head = 0;
    memset(out, 0, sizeof(double) * 9);
    out[0] = out[4] = out[8] = 1.0;
    while (head < count && count < 60) {
        const double* x = out + 9 * head;
        ++head;
        for (int g = 0; g < 2 && count < 60; ++g) {
            double y[9];
            for (int i = 0; i < 3; ++i)
                for (int j = 0; j < 3; ++j)
                    y[3 * i + j] = gen[g][3 * i] * x[j] + gen[g][3 * i + 1] * x[3 + j] + gen[g][3 * i + 2] * x[6 + j];
            int seen = 0;
            for (int k = 0; k < count && !seen; ++k) {
                double d = 0;
                for (int i = 0; i < 9; ++i) {
                    double e = fabs(out[9 * k + i] - y[i]);
                    if (e > d) d = e;
                }
                if (d < 1e-9) seen = 1;
            }
            if (!seen) { memcpy(out + 9 * count, y, sizeof(y)); ++count; }
        }
    }
}

/* The 60 operators in the reference's order, row-major 3x3 each. */
static void cim_icos60_ref_order(double* out /* 60*9 */)
{
    double canon[540];
    cim_icos60_canonical(canon);
    for (int k = 0; k < 60; ++k) memcpy(out + 9 * k, canon + 9 * cim_icos60_order[k], sizeof(double) * 9);
}

#endif
