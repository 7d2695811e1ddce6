// coarse_sep.hpp -- factors of the exact solve of the coarsest multigrid
// level (host code only; no HIP).
//
// Every level operator of the four systems is a sum of two Kronecker
// products of 1-D symmetric tridiagonal matrices,
//     A = R_t (x) S_p + P_t (x) Q_p,
// with t the direction that is transformed and p the one kept (R_t positive
// definite).  With the generalized eigenbasis P_t V = R_t V diag(lambda),
// V^T R_t V = I, the substitution z = (V (x) I) zh turns A z = b into one
// tridiagonal system per mode k along p:
//     (S_p + lambda_k Q_p) zh_k = bh_k,   bh = (V^T (x) I) b.
// The eigenproblem is reduced to a standard one through the Cholesky factor
// R_t = L L^T (C = L^-1 P_t L^-T, V = L^-T W) and solved by cyclic Jacobi
// rotations; each mode's system is factored once for the Thomas algorithm
// (no pivoting: every mode matrix here is definite).  A larger level is
// diagonalised in both directions instead (factor2): with Sp U = Qp U
// diag(mu), U^T Qp U = I as well, (V (x) U)^T A (V (x) U) = diag(lam_k +
// mu_j), and the solve is four dense transforms and a scaling, which several
// workgroups share.
#pragma once

#include <algorithm>
#include <cmath>
#include <vector>

namespace vdi {
namespace sep {

// Symmetric tridiagonal matrix: diagonal d, constant off-diagonal e.
struct Tri {
    std::vector<double> d;
    double e = 0;
};

// Cyclic Jacobi eigen-decomposition of the symmetric n x n matrix a
// (row-major, overwritten): eigenvalues w, eigenvectors as the columns of q
// (row-major).  Returns false if it has not converged after 64 sweeps.
inline bool jacobi(int n, std::vector<double> &a, std::vector<double> &w, std::vector<double> &q) {
    q.assign((size_t)n * n, 0.0);
    for (int i = 0; i < n; i++) q[(size_t)i * n + i] = 1;
    double norm = 0;
    for (double v : a) norm += v * v;
    bool ok = false;
    for (int sweep = 0; sweep < 64; sweep++) {
        double off = 0;
        for (int i = 0; i < n; i++)
            for (int j = i + 1; j < n; j++) off += a[(size_t)i * n + j] * a[(size_t)i * n + j];
        if (off <= 1e-30 * norm) { ok = true; break; }
        for (int p = 0; p < n; p++)
            for (int r = p + 1; r < n; r++) {
                const double apr = a[(size_t)p * n + r];
                if (apr == 0) continue;
                const double app = a[(size_t)p * n + p], arr = a[(size_t)r * n + r];
                // rotation angle that zeroes a[p][r]: tan(2 phi) = 2 apr / (arr - app)
                const double theta = (arr - app) / (2 * apr);
                const double t = (theta >= 0 ? 1.0 : -1.0) / (std::fabs(theta) + std::sqrt(theta * theta + 1));
                const double c = 1 / std::sqrt(t * t + 1), s = t * c;
                for (int k = 0; k < n; k++) {
                    const double akp = a[(size_t)k * n + p], akr = a[(size_t)k * n + r];
                    a[(size_t)k * n + p] = c * akp - s * akr;
                    a[(size_t)k * n + r] = s * akp + c * akr;
                }
                for (int k = 0; k < n; k++) {
                    const double apk = a[(size_t)p * n + k], ark = a[(size_t)r * n + k];
                    a[(size_t)p * n + k] = c * apk - s * ark;
                    a[(size_t)r * n + k] = s * apk + c * ark;
                }
                a[(size_t)p * n + r] = a[(size_t)r * n + p] = 0;
                for (int k = 0; k < n; k++) {
                    const double qkp = q[(size_t)k * n + p], qkr = q[(size_t)k * n + r];
                    q[(size_t)k * n + p] = c * qkp - s * qkr;
                    q[(size_t)k * n + r] = s * qkp + c * qkr;
                }
            }
    }
    w.resize(n);
    for (int i = 0; i < n; i++) w[i] = a[(size_t)i * n + i];
    return ok;
}

// Layout of the factors (also the kernel's LDS layout after the transformed
// right-hand side): V[t * T + k]; inv and cp, T rows of P + 1 (row k holds
// mode k's Thomas factors); then the T off-diagonals.
inline size_t factor_doubles(int T, int P) { return (size_t)T * T + 2 * (size_t)T * (P + 1) + T; }

// Generalized eigenbasis of the pencil (A, B), B positive definite: A V =
// B V diag(lam), V^T B V = I; V row-major, V[i * n + k] = component i of
// eigenvector k.  Returns false if the Cholesky factor or Jacobi fails.
inline bool geig(const Tri &A, const Tri &B, std::vector<double> &lam, std::vector<double> &V) {
    const int T = (int)A.d.size();
    std::vector<double> l(T), g(T, 0.0);
    for (int i = 0; i < T; i++) {
        if (i > 0) g[i] = B.e / l[i - 1];
        const double v = B.d[i] - g[i] * g[i];
        if (!(v > 0)) return false;
        l[i] = std::sqrt(v);
    }
    std::vector<double> x((size_t)T * T, 0.0), c((size_t)T * T, 0.0);
    for (int col = 0; col < T; col++)
        for (int i = 0; i < T; i++) {
            double v = (i == col ? A.d[i] : (i == col - 1 || i == col + 1 ? A.e : 0.0));
            if (i > 0) v -= g[i] * x[(size_t)(i - 1) * T + col];
            x[(size_t)i * T + col] = v / l[i];
        }
    for (int row = 0; row < T; row++)
        for (int i = 0; i < T; i++) {
            double v = x[(size_t)row * T + i];
            if (i > 0) v -= g[i] * c[(size_t)row * T + i - 1];
            c[(size_t)row * T + i] = v / l[i];
        }
    for (int i = 0; i < T; i++)
        for (int j = i + 1; j < T; j++) {
            const double v = 0.5 * (c[(size_t)i * T + j] + c[(size_t)j * T + i]);
            c[(size_t)i * T + j] = c[(size_t)j * T + i] = v;
        }
    std::vector<double> wv;
    if (!jacobi(T, c, lam, wv)) return false;
    V.assign((size_t)T * T, 0.0);
    for (int k = 0; k < T; k++)
        for (int i = T - 1; i >= 0; i--) {
            double v = wv[(size_t)i * T + k];
            if (i + 1 < T) v -= g[i + 1] * V[(size_t)(i + 1) * T + k];
            V[(size_t)i * T + k] = v / l[i];
        }
    return true;
}

// Layout of the factors of the fully diagonalised solve (both directions
// transformed): V (T x T), U (P x P) and its transpose, then 1 / (lam_k +
// mu_j) as T rows of P.
inline size_t factor2_doubles(int T, int P) { return (size_t)T * T + 2 * (size_t)P * P + (size_t)T * P; }

// Factors of A = Rt (x) Sp + Pt (x) Qp diagonalised in both directions:
// Pt V = Rt V diag(lam), Sp U = Qp U diag(mu) (Qp positive definite), so that
// (V (x) U)^T A (V (x) U) = diag(lam_k + mu_j).
inline bool factor2(const Tri &Pt, const Tri &Rt, const Tri &Sp, const Tri &Qp, std::vector<double> &out) {
    const int T = (int)Pt.d.size(), P = (int)Sp.d.size();
    std::vector<double> lam, V, mu, U;
    if (!geig(Pt, Rt, lam, V) || !geig(Sp, Qp, mu, U)) return false;
    out.assign(factor2_doubles(T, P), 0.0);
    double *o = out.data();
    std::copy(V.begin(), V.end(), o);
    o += (size_t)T * T;
    std::copy(U.begin(), U.end(), o);
    o += (size_t)P * P;
    for (int i = 0; i < P; i++)
        for (int j = 0; j < P; j++) o[(size_t)j * P + i] = U[(size_t)i * P + j];
    o += (size_t)P * P;
    for (int k = 0; k < T; k++)
        for (int j = 0; j < P; j++) {
            const double d = lam[k] + mu[j];
            if (!(std::fabs(d) > 0)) return false;
            o[(size_t)k * P + j] = 1 / d;
        }
    return true;
}

// The fully diagonalised solve with the factors of factor2, in the order of
// the kernels: rows times U, columns through V^T, the diagonal and V, rows
// times U^T.  b and z indexed [p + P * t].
inline void solve2(int T, int P, const std::vector<double> &f, const std::vector<double> &b, std::vector<double> &z) {
    const double *V = f.data(), *U = V + (size_t)T * T, *Ut = U + (size_t)P * P, *D = Ut + (size_t)P * P;
    std::vector<double> h((size_t)T * P, 0.0), g((size_t)T * P, 0.0);
    for (int t = 0; t < T; t++)
        for (int j = 0; j < P; j++) {
            double acc = 0;
            for (int p = 0; p < P; p++) acc += b[p + (size_t)P * t] * U[(size_t)p * P + j];
            h[j + (size_t)P * t] = acc;
        }
    for (int j = 0; j < P; j++) {
        std::vector<double> hk(T);
        for (int k = 0; k < T; k++) {
            double acc = 0;
            for (int t = 0; t < T; t++) acc += V[(size_t)t * T + k] * h[j + (size_t)P * t];
            hk[k] = acc * D[(size_t)k * P + j];
        }
        for (int t = 0; t < T; t++) {
            double acc = 0;
            for (int k = 0; k < T; k++) acc += V[(size_t)t * T + k] * hk[k];
            g[j + (size_t)P * t] = acc;
        }
    }
    z.assign((size_t)T * P, 0.0);
    for (int t = 0; t < T; t++)
        for (int p = 0; p < P; p++) {
            double acc = 0;
            for (int j = 0; j < P; j++) acc += g[j + (size_t)P * t] * Ut[(size_t)j * P + p];
            z[p + (size_t)P * t] = acc;
        }
}

// Factors of A = Rt (x) Sp + Pt (x) Qp (Rt of size T, Sp and Qp of size P).
// Returns false if the Cholesky factor, the eigen-decomposition or a Thomas
// pivot fails.
inline bool factor(const Tri &Pt, const Tri &Rt, const Tri &Sp, const Tri &Qp, std::vector<double> &out) {
    const int T = (int)Pt.d.size(), P = (int)Sp.d.size(), PP = P + 1;
    std::vector<double> lam, Vm;
    if (!geig(Pt, Rt, lam, Vm)) return false;
    out.assign(factor_doubles(T, P), 0.0);
    double *V = out.data(), *iv = V + (size_t)T * T, *cp = iv + (size_t)T * PP, *sub = cp + (size_t)T * PP;
    std::copy(Vm.begin(), Vm.end(), V);
    for (int k = 0; k < T; k++) {
        const double e = Sp.e + lam[k] * Qp.e;
        sub[k] = e;
        double cprev = 0;
        for (int p = 0; p < P; p++) {
            const double den = Sp.d[p] + lam[k] * Qp.d[p] - (p > 0 ? e * cprev : 0.0);
            if (!(std::fabs(den) > 0)) return false;
            iv[(size_t)k * PP + p] = 1 / den;
            cprev = e / den;
            cp[(size_t)k * PP + p] = cprev;
        }
    }
    return true;
}

// Pencils of the level operators; tdir 0 transforms along y (rows), 1 along
// x (columns).  The unknowns are indexed p + P t.

// Boundary types of the cell-centred operators and what they add to the
// diagonal factor of the boundary cell: Neumann 0, Dirichlet on the face 2,
// Dirichlet at the ghost centre 1.
inline double bc_factor(int bc) { return bc == 1 ? 2.0 : (bc == 2 ? 1.0 : 0.0); }

// 1-D second difference with boundary types lo / hi, times c, plus sigma.
inline Tri cc_line(int n, double c, double sigma, int lo, int hi) {
    Tri t;
    t.d.resize(n);
    for (int i = 0; i < n; i++)
        t.d[i] = sigma - c * ((i > 0 ? 1.0 : bc_factor(lo)) + (i < n - 1 ? 1.0 : bc_factor(hi)));
    t.e = c;
    return t;
}

// Cell-centred A = sigma I + s (cx Dx (x) I + cy I (x) Dy) on m x n cells.
inline void cc_pencil(double sigma, double s, int bcl, int bcr, int bcb, int bct, int m, int n, double cx, double cy,
                      int tdir, Tri &Pt, Tri &Rt, Tri &Sp, Tri &Qp) {
    const int T = tdir == 0 ? n : m, P = tdir == 0 ? m : n;
    Pt = tdir == 0 ? cc_line(n, s * cy, sigma, bcb, bct) : cc_line(m, s * cx, sigma, bcl, bcr);
    Sp = tdir == 0 ? cc_line(m, s * cx, 0.0, bcl, bcr) : cc_line(n, s * cy, 0.0, bcb, bct);
    Rt.d.assign(T, 1.0);
    Rt.e = 0;
    Qp.d.assign(P, 1.0);
    Qp.e = 0;
}

// Nodal Q1 operator on M x N nodes, row 0 Dirichlet (unknown rows 1..N-1):
// A = alpha Kx (x) My + beta Mx (x) Ky with the 1-D stiffness K (2, -1; 1 at
// a natural end) and mass M (4, 1; 2 at a natural end), the scales taken
// from the stencil's edge couplings.  Returns false if the stencil is not of
// that form.
inline bool fem_pencil(double fm, double fex, double fey, double fc, int M, int N, int tdir, Tri &Pt, Tri &Rt, Tri &Sp,
                       Tri &Qp) {
    const double beta = -(fex + 2 * fey) / 6, alpha = (2 * beta - fex) / 4;
    if (std::fabs(fm - 8 * (alpha + beta)) > 1e-12 * std::fabs(fm) ||
        std::fabs(fc + (alpha + beta)) > 1e-12 * std::fabs(fc))
        return false;
    Tri Kx, Mx, Ky, My;
    Kx.d.resize(M);
    Mx.d.resize(M);
    for (int i = 0; i < M; i++) {
        const bool end = i == 0 || i == M - 1;
        Kx.d[i] = end ? 1 : 2;
        Mx.d[i] = end ? 2 : 4;
    }
    Kx.e = -1;
    Mx.e = 1;
    Ky.d.resize(N - 1);
    My.d.resize(N - 1);
    for (int q = 0; q < N - 1; q++) {
        const bool top = q == N - 2;
        Ky.d[q] = top ? 1 : 2;
        My.d[q] = top ? 2 : 4;
    }
    Ky.e = -1;
    My.e = 1;
    auto scaled = [](Tri t, double f) {
        for (double &v : t.d) v *= f;
        t.e *= f;
        return t;
    };
    if (tdir == 0) {  // A = My (x) alpha Kx + Ky (x) beta Mx
        Rt = My;
        Sp = scaled(Kx, alpha);
        Pt = Ky;
        Qp = scaled(Mx, beta);
    } else {          // A = Mx (x) beta Ky + Kx (x) alpha My
        Rt = Mx;
        Sp = scaled(Ky, beta);
        Pt = Kx;
        Qp = scaled(My, alpha);
    }
    return true;
}

// The exact solve with the factors, in the kernel's order of operations
// (host reference for checks): b and z indexed [p + P * t].
inline void solve(int T, int P, const std::vector<double> &f, const std::vector<double> &b, std::vector<double> &z) {
    const int PP = P + 1;
    const double *V = f.data(), *iv = V + (size_t)T * T, *cp = iv + (size_t)T * PP, *sub = cp + (size_t)T * PP;
    std::vector<double> h((size_t)T * PP, 0.0);
    for (int k = 0; k < T; k++)
        for (int p = 0; p < P; p++) {
            double acc = 0;
            for (int t = 0; t < T; t++) acc += V[(size_t)t * T + k] * b[p + (size_t)P * t];
            h[(size_t)k * PP + p] = acc;
        }
    for (int k = 0; k < T; k++) {
        double *hk = h.data() + (size_t)k * PP, y = 0;
        for (int p = 0; p < P; p++) hk[p] = y = (hk[p] - sub[k] * y) * iv[(size_t)k * PP + p];
        for (int p = P - 2; p >= 0; p--) hk[p] = y = hk[p] - cp[(size_t)k * PP + p] * y;
    }
    z.assign((size_t)T * P, 0.0);
    for (int t = 0; t < T; t++)
        for (int p = 0; p < P; p++) {
            double acc = 0;
            for (int k = 0; k < T; k++) acc += V[(size_t)t * T + k] * h[(size_t)k * PP + p];
            z[p + (size_t)P * t] = acc;
        }
}

}  // namespace sep
}  // namespace vdi
