// pg_herm.h -- host-side FP64 eigen-decomposition of a small Hermitian
// matrix (the min-max time-segmentation interpolator's Gram matrix).
#pragma once

#include <algorithm>
#include <cmath>
#include <complex>
#include <vector>

namespace pg {

// Eigen-decomposition of a Hermitian n x n matrix (row-major) by cyclic
// Jacobi rotations in FP64: A = V diag(w) V^H.  Each rotation first turns
// a_pq real with a phase on column q, then applies the real symmetric
// rotation (Numerical Recipes' jacobi: theta = (a_qq - a_pp)/(2 a_pq)).
inline void herm_eig(int n, std::vector<std::complex<double>> A, std::vector<double>& w,
                     std::vector<std::complex<double>>& V)
{
    using cd = std::complex<double>;
    V.assign((size_t)n * n, cd(0, 0));
    for (int k = 0; k < n; ++k) V[(size_t)k * n + k] = 1.0;
    double scale = 0;
    for (const cd& a : A) scale = std::max(scale, std::abs(a));
    for (int sweep = 0; sweep < 100; ++sweep) {
        double off = 0;
        for (int p = 0; p < n; ++p)
            for (int q = p + 1; q < n; ++q) off += std::norm(A[(size_t)p * n + q]);
        if (std::sqrt(off) <= 1e-17 * scale * n) break;
        for (int p = 0; p < n; ++p)
            for (int q = p + 1; q < n; ++q) {
                const cd apq = A[(size_t)p * n + q];
                const double r = std::abs(apq);
                if (r <= 1e-300) continue;
                const cd e = apq / r;
                const double app = A[(size_t)p * n + p].real(), aqq = A[(size_t)q * n + q].real();
                const double th = (aqq - app) / (2.0 * r);
                const double t = (th >= 0 ? 1.0 : -1.0) / (std::fabs(th) + std::sqrt(th * th + 1.0));
                const double c = 1.0 / std::sqrt(t * t + 1.0), sn = t * c;
                // U = diag(1, conj(e)) R on (p, q): U_pp = c, U_pq = s,
                // U_qp = -s conj(e), U_qq = c conj(e); A <- U^H A U, V <- V U
                const cd ce = std::conj(e);
                for (int k = 0; k < n; ++k) {
                    const cd akp = A[(size_t)k * n + p], akq = A[(size_t)k * n + q];
                    A[(size_t)k * n + p] = c * akp - sn * ce * akq;
                    A[(size_t)k * n + q] = sn * akp + c * ce * akq;
                    const cd vkp = V[(size_t)k * n + p], vkq = V[(size_t)k * n + q];
                    V[(size_t)k * n + p] = c * vkp - sn * ce * vkq;
                    V[(size_t)k * n + q] = sn * vkp + c * ce * vkq;
                }
                for (int k = 0; k < n; ++k) {
                    const cd apk = A[(size_t)p * n + k], aqk = A[(size_t)q * n + k];
                    A[(size_t)p * n + k] = c * apk - sn * e * aqk;
                    A[(size_t)q * n + k] = sn * apk + c * e * aqk;
                }
                A[(size_t)p * n + q] = A[(size_t)q * n + p] = 0.0;
            }
    }
    w.resize(n);
    for (int k = 0; k < n; ++k) w[k] = A[(size_t)k * n + k].real();
}

}  // namespace pg
