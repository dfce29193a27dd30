// jg_csr.h -- Jet's compressed linear system on the device: FdmCompressedBlas3
// (src/jet/fdm_linear_system3.cpp:136-207) over a MatrixCsrD whose rows are
// the fluid cells only (include/jet/matrix_csr.h:96-478).
//
// Layout: CSR exactly as MatrixCsr stores it -- values, row pointers
// (rows + 1) and column indices, a row's entries in ascending column order
// (MatrixCsr::addRow sorts them, include/jet/detail/matrix_csr-inl.h:253-275)
// -- with 32-bit indices on the device.  One thread per row; a row's sum
// runs over its entries in storage order, as Jet's mvm / residual do, so
// products and residuals are bitwise Jet's.
#pragma once

#include "jg_device.h"

namespace jg {

struct CsrDev {
    const int* rp;      // rows + 1 row pointers
    const int* ci;      // column indices
    const double* val;  // non-zeros
    int nnz;            // entries: a row's range is clamped to it
};

// Entry range of row i, clamped to [0, nnz] so that no row pointer can send
// a thread past the arrays.
__device__ __forceinline__ void csr_row(const CsrDev& m, idx_t i, int& b, int& e) {
    b = min(max(m.rp[i], 0), m.nnz);
    e = min(max(m.rp[i + 1], b), m.nnz);
}

// FdmCompressedBlas3::mvm (:160-179): y_i = sum_jj val[jj] x[ci[jj]].
__global__ __launch_bounds__(NT) void k_csr_mvm(CsrDev m, const double* __restrict__ x,
                                                double* __restrict__ y, idx_t n) {
#pragma clang fp contract(off)
    for (idx_t i = blockIdx.x * (idx_t)NT + threadIdx.x; i < n; i += (idx_t)gridDim.x * NT) {
        int jb, je;
        csr_row(m, i, jb, je);
        double sum = 0.0;
        for (int jj = jb; jj < je; ++jj) sum += m.val[jj] * x[m.ci[jj]];
        y[i] = sum;
    }
}

// FdmCompressedBlas3::residual (:181-199): r_i = b_i - (A x)_i.  DOT: partial
// r.r.  REFRESH: only on iterations where the control block asks for a true
// residual (Jet's every-50 / sigma-growth rule, include/jet/detail/cg-inl.h:59-66).
template <bool DOT, bool REFRESH>
__global__ __launch_bounds__(NT) void k_csr_residual(CsrDev m, const double* __restrict__ x,
                                                     const double* __restrict__ b,
                                                     double* __restrict__ r, idx_t n,
                                                     double* __restrict__ partial,
                                                     const Ctl* __restrict__ ctl) {
#pragma clang fp contract(off)
    if (REFRESH && (ctl->done || !ctl->refresh)) return;
    double acc = 0.0;
    for (idx_t i = blockIdx.x * (idx_t)NT + threadIdx.x; i < n; i += (idx_t)gridDim.x * NT) {
        int jb, je;
        csr_row(m, i, jb, je);
        double sum = 0.0;
        for (int jj = jb; jj < je; ++jj) sum += m.val[jj] * x[m.ci[jj]];
        const double rv = b[i] - sum;
        r[i] = rv;
        if (DOT) acc += rv * rv;
    }
    if (DOT) block_partial<NT>(acc, partial, (idx_t)blockIdx.x);
}

// The search direction and its product in one pass: d = beta d + z (Jet's
// axpy(beta, d, s, d), cg-inl.h:85), q = A d (mvm, :50), and the partial d.q.
// A row needs d at its columns, which this pass also produces, so the new d
// is recomputed there from z and the old d (the same expression, so the
// same value).
__global__ __launch_bounds__(NT) void k_csr_dir_spmv(CsrDev m, const double* __restrict__ z,
                                                     const double* __restrict__ dold,
                                                     double* __restrict__ dnew,
                                                     double* __restrict__ q, idx_t n,
                                                     double* __restrict__ partial,
                                                     const Ctl* __restrict__ ctl) {
#pragma clang fp contract(off)
    if (ctl->done) return;
    const double beta = ctl->beta;
    double acc = 0.0;
    for (idx_t i = blockIdx.x * (idx_t)NT + threadIdx.x; i < n; i += (idx_t)gridDim.x * NT) {
        int jb, je;
        csr_row(m, i, jb, je);
        double sum = 0.0;
        for (int jj = jb; jj < je; ++jj) {
            const int j = m.ci[jj];
            sum += m.val[jj] * (beta * dold[j] + z[j]);
        }
        const double di = beta * dold[i] + z[i];
        dnew[i] = di;
        q[i] = sum;
        acc += di * sum;
    }
    block_partial<NT>(acc, partial, (idx_t)blockIdx.x);
}

// ELL copy of a compressed system whose rows have at most kEllMaxW entries
// (a pressure row has at most 7): entry k of row i at k * rows + i, so one
// thread per row reads its entries with fully coalesced loads; len[i]
// entries are used, in the CSR order, so every sum is the CSR kernels'.
constexpr int kEllMaxW = 8;

struct EllDev {
    const double* val;
    const int* ci;
    const int8_t* len;
    idx_t rows;
};

__global__ __launch_bounds__(NT) void k_csr_to_ell(CsrDev m, idx_t rows, double* __restrict__ val,
                                                   int* __restrict__ ci, int8_t* __restrict__ len,
                                                   int w) {
    for (idx_t i = blockIdx.x * (idx_t)NT + threadIdx.x; i < rows; i += (idx_t)gridDim.x * NT) {
        int jb, je;
        csr_row(m, i, jb, je);
        const int l = min(je - jb, w);
        len[i] = (int8_t)l;
        for (int k = 0; k < w; ++k) {
            val[k * rows + i] = k < l ? m.val[jb + k] : 0.0;
            ci[k * rows + i] = k < l ? m.ci[jb + k] : (int)i;
        }
    }
}

// k_csr_residual on the ELL copy.
template <bool DOT, bool REFRESH>
__global__ __launch_bounds__(NT) void k_ell_residual(EllDev m, const double* __restrict__ x,
                                                     const double* __restrict__ b,
                                                     double* __restrict__ r,
                                                     double* __restrict__ partial,
                                                     const Ctl* __restrict__ ctl) {
#pragma clang fp contract(off)
    if (REFRESH && (ctl->done || !ctl->refresh)) return;
    double acc = 0.0;
    for (idx_t i = blockIdx.x * (idx_t)NT + threadIdx.x; i < m.rows; i += (idx_t)gridDim.x * NT) {
        const int l = m.len[i];
        double sum = 0.0;
        for (int k = 0; k < l; ++k) sum += m.val[k * m.rows + i] * x[m.ci[k * m.rows + i]];
        const double rv = b[i] - sum;
        r[i] = rv;
        if (DOT) acc += rv * rv;
    }
    if (DOT) block_partial<NT>(acc, partial, (idx_t)blockIdx.x);
}

// The direction update alone on the compressed rows: d = beta d_old + z
// (k_csr_dir_spmv's expression).
__global__ __launch_bounds__(NT) void k_ell_dir(const double* __restrict__ z,
                                               const double* __restrict__ dold,
                                               double* __restrict__ dnew, idx_t n,
                                               const Ctl* __restrict__ ctl) {
#pragma clang fp contract(off)
    if (ctl->done) return;
    const double beta = ctl->beta;
    for (idx_t i = blockIdx.x * (idx_t)NT + threadIdx.x; i < n; i += (idx_t)gridDim.x * NT)
        dnew[i] = beta * dold[i] + z[i];
}

// q = A d on the ELL copy, with one partial of d.q per block (after
// k_ell_dir: the values of k_csr_dir_spmv, gathering one vector per entry).
__global__ __launch_bounds__(NT) void k_ell_spmv_dot(EllDev m, const double* __restrict__ d,
                                                    double* __restrict__ q,
                                                    double* __restrict__ partial,
                                                    const Ctl* __restrict__ ctl) {
#pragma clang fp contract(off)
    if (ctl->done) return;
    double acc = 0.0;
    for (idx_t i = blockIdx.x * (idx_t)NT + threadIdx.x; i < m.rows; i += (idx_t)gridDim.x * NT) {
        const int l = m.len[i];
        double sum = 0.0;
        for (int k = 0; k < l; ++k) sum += m.val[k * m.rows + i] * d[m.ci[k * m.rows + i]];
        q[i] = sum;
        acc += d[i] * sum;
    }
    block_partial<NT>(acc, partial, (idx_t)blockIdx.x);
}

}  // namespace jg
