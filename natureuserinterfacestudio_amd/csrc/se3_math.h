// FP64 algebra of the compensated ICP mode: the 6x6 LDL^T solve of the normal
// equations and the SE(3) exponential map that applies the increment.  Host and
// device compile the same code with -ffp-contract=off and use only correctly
// rounded IEEE operations (+, -, *, /), so both produce identical bits; the sine
// and cosine terms of the exponential are power series evaluated in a fixed
// order instead of library calls whose last bit differs between libraries.
#pragma once

#include "det_math.h"

namespace nuif_se3 {

// A = L D L^T in place (unit lower L below the diagonal, D on it); det receives
// the product of the pivots.  Returns false at the first pivot that is not
// positive (A not positive definite).
NUIF_HD bool ldlt6(double A[36], double* det) {
    double d = 1.0;
    for (int j = 0; j < 6; ++j) {
        double Dj = A[j * 6 + j];
        for (int k = 0; k < j; ++k) Dj -= A[j * 6 + k] * A[j * 6 + k] * A[k * 6 + k];
        A[j * 6 + j] = Dj;
        d *= Dj;
        if (!(Dj > 0.0)) {
            *det = d;
            return false;
        }
        for (int i = j + 1; i < 6; ++i) {
            double v = A[i * 6 + j];
            for (int k = 0; k < j; ++k) v -= A[i * 6 + k] * A[j * 6 + k] * A[k * 6 + k];
            A[i * 6 + j] = v / Dj;
        }
    }
    *det = d;
    return true;
}

// x = A^-1 b from the factor of ldlt6
NUIF_HD void ldlt6_solve(const double F[36], const double b[6], double x[6]) {
    double y[6];
    for (int i = 0; i < 6; ++i) {
        double s = b[i];
        for (int k = 0; k < i; ++k) s -= F[i * 6 + k] * y[k];
        y[i] = s;
    }
    for (int i = 0; i < 6; ++i) y[i] /= F[i * 6 + i];
    for (int i = 5; i >= 0; --i) {
        double s = y[i];
        for (int k = i + 1; k < 6; ++k) s -= F[k * 6 + i] * x[k];
        x[i] = s;
    }
}

// Largest rotation angle (radians) of an increment the series below evaluate to
// full double precision; a Gauss-Newton step beyond it is a tracking failure.
constexpr double kMaxStepAngle = 4.0;

// The coefficients of exp([w]x) = I + A [w]x + B [w]x^2 and of its left Jacobian
// V = I + B [w]x + C [w]x^2 as power series in th2 = |w|^2:
//   A = sum (-th2)^k / (2k+1)!,  B = sum (-th2)^k / (2k+2)!,  C = sum (-th2)^k / (2k+3)!
// (sin th / th, (1 - cos th) / th^2, (th - sin th) / th^3), 22 terms each, Horner
// order; for th <= kMaxStepAngle the first omitted term is below 1e-19 of the sum.
NUIF_HD void se3_coefficients(double th2, double* A, double* B, double* C) {
    // 1 / m! for m = 0..45, each the correctly rounded double
    constexpr double kInvFact[46] = {
        1.0, 1.0, 0.5, 0.16666666666666666,
        0.041666666666666664, 0.008333333333333333, 0.001388888888888889, 0.0001984126984126984,
        2.48015873015873e-05, 2.7557319223985893e-06, 2.755731922398589e-07, 2.505210838544172e-08,
        2.08767569878681e-09, 1.6059043836821613e-10, 1.1470745597729725e-11, 7.647163731819816e-13,
        4.779477332387385e-14, 2.8114572543455206e-15, 1.5619206968586225e-16, 8.22063524662433e-18,
        4.110317623312165e-19, 1.9572941063391263e-20, 8.896791392450574e-22, 3.868170170630684e-23,
        1.6117375710961184e-24, 6.446950284384474e-26, 2.4795962632247976e-27, 9.183689863795546e-29,
        3.279889237069838e-30, 1.1309962886447716e-31, 3.7699876288159054e-33, 1.216125041553518e-34,
        3.8003907548547434e-36, 1.151633562077195e-37, 3.387157535521162e-39, 9.67759295863189e-41,
        2.6882202662866363e-42, 7.265460179153071e-44, 1.911963205040282e-45, 4.902469756513544e-47,
        1.2256174391283858e-48, 2.9893108271424046e-50, 7.117406731291439e-52, 1.6552108677421951e-53,
        3.7618428812322616e-55, 8.359650847182804e-57,
    };
    constexpr int kSeries = 22;
    double a = 0.0, b = 0.0, c = 0.0;
    for (int k = kSeries - 1; k >= 0; --k) {
        a = a * -th2 + kInvFact[2 * k + 1];
        b = b * -th2 + kInvFact[2 * k + 2];
        c = c * -th2 + kInvFact[2 * k + 3];
    }
    *A = a;
    *B = b;
    *C = c;
}

// The twist (w, u) as a rigid motion p' = R p + t: R = exp([w]x), t = V u.
// Returns false when |w| exceeds kMaxStepAngle.
NUIF_HD bool se3_exp(const double w[3], const double u[3], double R[9], double t[3]) {
    const double th2 = w[0] * w[0] + w[1] * w[1] + w[2] * w[2];
    if (!(th2 <= kMaxStepAngle * kMaxStepAngle)) return false;
    double A, B, C;
    se3_coefficients(th2, &A, &B, &C);
    const double K[9] = {0.0, -w[2], w[1], w[2], 0.0, -w[0], -w[1], w[0], 0.0};
    double K2[9];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
            K2[3 * i + j] = K[3 * i] * K[j] + K[3 * i + 1] * K[3 + j] + K[3 * i + 2] * K[6 + j];
    double V[9];
    for (int k = 0; k < 9; ++k) {
        const double I = (k % 4 == 0) ? 1.0 : 0.0;
        R[k] = I + A * K[k] + B * K2[k];
        V[k] = I + B * K[k] + C * K2[k];
    }
    for (int i = 0; i < 3; ++i) t[i] = V[3 * i] * u[0] + V[3 * i + 1] * u[1] + V[3 * i + 2] * u[2];
    return true;
}

// (R, t) <- (Ri R, Ri t + ti): the increment applied on the left, in world coordinates.
NUIF_HD void se3_compose_left(const double Ri[9], const double ti[3], double R[9], double t[3]) {
    double Rn[9], tn[3];
    for (int i = 0; i < 3; ++i) {
        for (int j = 0; j < 3; ++j) Rn[3 * i + j] = Ri[3 * i] * R[j] + Ri[3 * i + 1] * R[3 + j] + Ri[3 * i + 2] * R[6 + j];
        tn[i] = Ri[3 * i] * t[0] + Ri[3 * i + 1] * t[1] + Ri[3 * i + 2] * t[2] + ti[i];
    }
    for (int k = 0; k < 9; ++k) R[k] = Rn[k];
    for (int k = 0; k < 3; ++k) t[k] = tn[k];
}

// One Gauss-Newton step of the compensated mode from the 29 FP64 sums of an
// iteration (27 unique entries of [A | b] in row order, r^2, count):
//   A = sums / count, b likewise, undamped; LDL^T solve; the increment x =
//   (x_rot, x_t) is applied as exp of the twist (-x_rot, x_t) on the left.
// ws: 60 doubles of workspace (the system, the solution and the increment; the
// device passes shared memory to keep registers free).
// Returns 0: applied, 1: converged (|x|^2 < step2_stop, not applied), 2: failure
// (not positive definite, |det| < 1e-15, or a rotation step beyond kMaxStepAngle).
NUIF_HD int gn_step(const double sums[29], double step2_stop, double R[9], double t[3], double* ws) {
    double* A = ws;        // 36
    double* b = ws + 36;   // 6
    double* x = ws + 42;   // 6
    double* Ri = ws + 48;  // 9
    double* ti = ws + 57;  // 3
    const double cnt = sums[28];
    for (int i = 0, s = 0; i < 6; ++i)
        for (int j = i; j < 7; ++j, ++s) {
            const double v = sums[s] / cnt;
            if (j == 6)
                b[i] = v;
            else
                A[j * 6 + i] = A[i * 6 + j] = v;
        }
    double det;
    const bool spd = ldlt6(A, &det);
    if (!spd || fabs(det) < 1e-15 || det != det) return 2;
    ldlt6_solve(A, b, x);
    double n2 = 0.0;
    for (int k = 0; k < 6; ++k) n2 += x[k] * x[k];
    if (n2 < step2_stop) return 1;
    const double w[3] = {-x[0], -x[1], -x[2]}, u[3] = {x[3], x[4], x[5]};
    if (!se3_exp(w, u, Ri, ti)) return 2;
    se3_compose_left(Ri, ti, R, t);
    return 0;
}
NUIF_HD int gn_step(const double sums[29], double step2_stop, double R[9], double t[3]) {
    double ws[60];
    return gn_step(sums, step2_stop, R, t, ws);
}

// Convergence threshold of gn_step: a step below 1e-6 (radians and metres) stops the level.
constexpr double kGnStep2 = 1e-12;

// Pixels per work item of the compensated ICP order (256 lanes, 2 pixels each): the
// order of its sums depends on it.
constexpr int kCompTilePixels = 512;

}  // namespace nuif_se3
