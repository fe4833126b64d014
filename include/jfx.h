/*
 * jfx.h -- C ABI of libjfx.so, the MI355X (gfx950) kernels behind the
 * spectral-Galerkin hot path of jaxfun_amd.
 *
 * Conventions
 *   - All pointers are device pointers unless a comment says "host".
 *   - Arrays are C-order.  A "batch along an axis" is viewed as (P, n, Q):
 *     P = product of the leading dims, Q = product of the trailing dims.
 *   - Every entry point is stream-ordered on `stream` (a hipStream_t; NULL =
 *     the legacy default stream) and returns 0 or a negative status:
 *       -1 HIP error, -2 bad/unsupported sizes, -3 singular pivot,
 *       -4 bad argument.  jfx_last_error() returns the message (per thread).
 *   - No entry point allocates device memory unless its comment says so.
 *
 * Each declaration cites the reference interface (paths relative to the
 * reference's src/jaxfun) that it replaces on the device.
 */
#ifndef JFX_H
#define JFX_H

#include <hip/hip_runtime.h>

#ifdef __cplusplus
extern "C" {
#endif

int jfx_version(void);
const char* jfx_last_error(void);
int jfx_device_sync(void);

/* ---- quadrature.  family 0 = Legendre, 1 = Chebyshev.  Nodes ascending.
 * Replaces Legendre.quad_points_and_weights (galerkin/Legendre.py:126-137 ->
 * utils/fastgl.py:561-567) and Chebyshev.quad_points_and_weights
 * (galerkin/Chebyshev.py:149-170). */
int jfx_gauss_nodes(int family, int N, double* x, double* w, hipStream_t stream);

/* ---- Vandermonde: V[q, n] = scale * d^k/dx^k psi_n(x[q]), V is (Q, N), k <= 4.
 * Replaces OrthogonalSpace.vandermonde / evaluate_basis_derivative
 * (galerkin/orthogonal.py:131-141, 168-178) over eval_basis_functions
 * (galerkin/Legendre.py:163-183, galerkin/Chebyshev.py:200-223). */
int jfx_vandermonde_f64(int family, const double* x, int Q, int N, int k, double scale, double* V,
                        hipStream_t stream);

/* Vandermonde of a basis given by its three-term recurrence, V (Q, N),
 * V[q, n] = scale * d^k/dx^k Q_n(x[q]), k <= 4, with Q_0 = 1 and
 * Q_n = r0[n] ((x - r1[n]) Q_{n-1} - r2[n] Q_{n-2}), rec = [r0 | r1 | r2]
 * (3 N doubles, device).  Replaces Jacobi.eval_basis_functions
 * (galerkin/Jacobi.py:165-202) and its nested-jacfwd derivatives
 * (galerkin/orthogonal.py:168-178) for Jacobi, ChebyshevU and Ultraspherical
 * (galerkin/ChebyshevU.py:134-157, galerkin/Ultraspherical.py:12-101). */
int jfx_vandermonde_rec_f64(const double* x, int Q, int N, int k, const double* rec, double scale, double* V,
                            hipStream_t stream);

/* Fourier basis (complex, re/im interleaved): element (q, j) of
 * (i k_j s)^d exp(i k_j x_q) * post for the N-mode space (FFT wavenumber
 * order; Nyquist mode zeroed for odd d and even N) at x_q = 2 pi q / Q
 * (grid != 0, exact phase reduction) or x[q]; out (Q, N), or its conjugate
 * transpose (N, Q) when adjoint != 0.  Replaces
 * Fourier.eval_basis_functions / evaluate_basis_derivative and the FFTs of
 * Fourier.backward / forward / scalar_product (galerkin/Fourier.py:92-180),
 * which become contractions with these matrices. */
int jfx_fourier_basis(const double* x, int Q, int N, int d, double s, int grid, int adjoint, double post,
                      double* out, hipStream_t stream);

/* out[c, r] = V[r, c] * rowscale[r] * colscale[c] * s for an (R, C) matrix
 * (either scale pointer may be NULL).  Builds the quadrature-weighted
 * transposed Vandermonde (P^T * w) of OrthogonalSpace.scalar_product
 * (galerkin/orthogonal.py:264-277) and of inner_bilinear's quadrature branch
 * (galerkin/inner.py:887-900). */
int jfx_scale_transpose(const double* V, int R, int C, const double* rowscale, const double* colscale,
                        double s, double* out, hipStream_t stream);

/* ---- per-axis contraction Y[p, m, q] = sum_k A[m, k] X[p, k, q];
 * A is (M, K) row-major, X is (P, K, Q), Y is (P, M, Q).
 * Replaces DiaMatrix/Matrix.matvec(axis=i) (la/diamatrix.py:227-349,
 * la/matrix.py:89-118) as used by TPMatrix._matmul_array
 * (la/tpmatrix.py:152-156) and by TensorProductSpace.forward/backward/
 * scalar_product (galerkin/tensorproductspace.py:330-417). */
int jfx_contract_f64(const double* A, int M, int K, const double* X, double* Y, long long P,
                     long long Q, hipStream_t stream);
int jfx_contract_f32(const float* A, int M, int K, const float* X, float* Y, long long P,
                     long long Q, hipStream_t stream);

/* Split an FP64 (M, K) matrix into `nslices` bf16 slices stored [s][M][K].
 * The contraction kernels read slices with a row stride Kp = 32*ceil(K/32):
 * split a zero-padded (M, Kp) copy of A to obtain that layout.
 * mode 0: float split of fp32(A) (for jfx_contract_bf16x);
 * mode 1: fixed-point digits with power-of-two row scales rowscale[M]
 *         (for jfx_contract_ozaki).  Same reference interface as above. */
int jfx_split_matrix(const double* A, int M, int K, int mode, int nslices, unsigned short* slices,
                     double* rowscale, hipStream_t stream);

/* BF16 float-split contraction on FP32 data: nslices 2 -> 3 products,
 * 3 -> 6 products, FP32 accumulation.  Same reference interface as above. */
int jfx_contract_bf16x(const unsigned short* Aslices, int nslices, int M, int K, const float* X,
                       float* Y, long long P, long long Q, hipStream_t stream);

/* Fixed-point-slice contraction on FP64 data: nslices in {3, 4, 5} 8-bit
 * digit slices per operand, bf16 MFMA with exact FP32 accumulation
 * (requires nslices * K <= 1024), levels combined in FP64.
 * Same reference interface as above. */
int jfx_contract_ozaki(const unsigned short* Aslices, const double* Arowscale, int nslices, int M,
                       int K, const double* X, double* Y, long long P, long long Q,
                       hipStream_t stream);

/* Split an FP64 (M, K) matrix into nslices (5, 6 or 7) balanced int8 digit
 * slices [s][M][Kp] (Kp = 64 ceil(K / 64), zero padded), digits in [-64, 64]:
 * A[m, k] = 2^rowexp[m] sum_s slices[s][m][k] 2^(-6-7s) up to 2^(-6-7(nslices-1))
 * of the row maximum (|A[m, :]| < 2^rowexp[m]).  For jfx_contract_i8.
 * Same reference interface as the contractions above. */
int jfx_split_matrix_i8(const double* A, int M, int K, int nslices, signed char* slices, int* rowexp,
                        hipStream_t stream);

/* FP64 contraction on int8 MFMA (M, K <= 128; Q == 1 with K even, or
 * Q % 64 == 0; X 16-byte aligned): X is cut per column into the same
 * balanced digits, the nslices (nslices + 1) / 2 digit products of levels
 * s + t < nslices accumulate exactly in int32 and are combined in FP64.
 * Error ~ 2^(-6-7(nslices-1)) of the row and column maxima.  Same reference
 * interface as the contractions above (galerkin/tensorproductspace.py:330-417
 * for the 3D transforms). */
int jfx_contract_i8(const signed char* Aslices, const int* Arowexp, int nslices, int M, int K, const double* X,
                    double* Y, long long P, long long Q, hipStream_t stream);

/* Fused two-axis FP64 contraction of 2D slabs on int8 MFMA: Y[b] = A0 X[b] A1^T
 * (X: B x K0 x K1, Y: B x M0 x M1, FP64, all dims <= 128, K1 even, X and Y
 * 16-byte aligned), with the operators as jfx_split_matrix_i8 digit slices
 * (nslices 5 or 6).  One HBM pass for the last two axes of a transform
 * (galerkin/tensorproductspace.py:330-417) instead of two jfx_contract_i8
 * calls; same error bound per contraction. */
int jfx_plane_i8(const signed char* A0slices, const int* A0rowexp, const signed char* A1slices, const int* A1rowexp,
                 int nslices, const double* X, double* Y, long long B, int M0, int K0, int M1, int K1,
                 hipStream_t stream);

/* Fused two-axis contraction of 2D slabs: Y[b] = alpha A0 X[b] A1^T (X: B x K0
 * x K1, Y: B x M0 x M1, all dims <= 128), optionally followed by Y /= E (epi 1)
 * or Y *= E (epi 2) with E (M0 x M1); accumulate = 1 adds the result to the
 * previous Y (Y must not alias X).  prec 0: FP64 operators A0d, A1d; prec 1:
 * 3-term bf16 slices A0s, A1s ([3][M][Kp], Kp = 32*ceil(K/32)) on FP32 data.
 * Replaces two successive per-axis applications of TensorProductSpace
 * transforms (galerkin/tensorproductspace.py:330-417), the 2D
 * fast-diagonalisation solve (la/tpmatrix.py:545-587, with the eigenvalue
 * division as the epilogue) and, with alpha and accumulate, the scaled sum
 * of a TPMatrices product (la/tpmatrix.py:158-166, 291-316). */
int jfx_plane_contract(int prec, const double* A0d, const double* A1d, const unsigned short* A0s,
                       const unsigned short* A1s, const void* X, void* Y, const void* E, int epi, double alpha,
                       int accumulate, long long B, int M0, int K0, int M1, int K1, hipStream_t stream);

/* 2D fast-diagonalisation solve of n x n slabs in one launch:
 *     Y[b] = V0 ((W0 X[b] W1^T) / E) V1^T        (X, Y: B x n x n, E: n x n),
 * the intermediate kept in LDS.  prec 0: FP64 operators, n <= 64; prec 1:
 * 3-term bf16 slices ([3][n][Kp]) on FP32 data, n <= 128.  Replaces
 * TPMatricesLUFactors.solve for two axes (la/tpmatrix.py:545-587). */
int jfx_plane_fd(int prec, const double* W0d, const double* W1d, const double* V0d, const double* V1d,
                 const unsigned short* W0s, const unsigned short* W1s, const unsigned short* V0s,
                 const unsigned short* V1s, const void* X, void* Y, const void* E, long long B, int n,
                 hipStream_t stream);

/* ---- banded LU without pivoting.  band is (n, p+q+1) "by row":
 * band[i*(p+q+1) + o + p] = A[i, i+o].  status_dev (device int) receives 0,
 * or 1 + the row of the first zero/non-finite pivot.
 * Replaces DiaMatrix.lu_factor no-pivot kernel (la/diamatrix.py:394-504,
 * 1936-1973). */
int jfx_band_lu_factor(const double* band, int n, int p, int q, double* lu, int* status_dev,
                       hipStream_t stream);
/* In-place solve along the middle axis of X (P, n, Q).
 * Replaces LUFactors.solve (la/diamatrix.py:1880-1925) with
 * _forward_elimination/_backward_substitution (la/diamatrix.py:2079-2195). */
int jfx_band_lu_solve(const double* lu, int n, int p, int q, double* X, long long P, long long Q,
                      hipStream_t stream);
/* As jfx_band_lu_solve with the kernel chosen by mode: 1 = one lane per line
 * (sequential sweep), 2 = wavefronts per line (the rows split into chunks over
 * 64 lanes, or 512 lanes for at most 256 lines of n >= 4096; each chunk's
 * affine map of the recurrence window is composed by a wave scan; p, q <= 4,
 * else JFX_EARG), 0 = automatic (mode 2 for p, q <= 4, n >= 32 and at most
 * 2048 lines, mode 1 otherwise; jfx_band_lu_solve is mode 0).  Replaces the
 * same LUFactors.solve (la/diamatrix.py:1880-1925, 2079-2195) for a single
 * system or a few lines: DiaMatrix.solve on a vector
 * (la/diamatrix.py:592-770). */
int jfx_band_lu_solve_mode(const double* lu, int n, int p, int q, double* X, long long P, long long Q, int mode,
                           hipStream_t stream);
/* The same solve for few long lines (p, q <= 4, n >= 8192, lines x chunks
 * <= 1024) with each line's rows split over several 8-wave workgroups (chunks
 * of >= 4096 rows, at most 64 per line), one launch per direction: each
 * workgroup takes the recurrence window entering its rows from its
 * predecessor through device flags (release / acquire at agent scope;
 * bounded waits, a chunk that gives up propagates NaN instead of hanging).
 * work holds jfx_band_lu_solve_grid_work_bytes(...) bytes (0: the shape does
 * not suit the grid solve); calls sharing a work buffer must not overlap.
 * Same reference interface as jfx_band_lu_solve. */
long long jfx_band_lu_solve_grid_work_bytes(int n, int p, int q, long long P, long long Q);
int jfx_band_lu_solve_grid(const double* lu, int n, int p, int q, double* X, long long P, long long Q, void* work,
                           long long work_bytes, hipStream_t stream);

/* Dense LU with partial pivoting, in place, of a row-major n x n matrix A
 * (cplx = 0: FP64, cplx = 1: complex128 as re/im pairs), n <= 16384:
 * blocked over 32-column panels (16 complex; one workgroup per panel with
 * the panel's rows in registers while at most 4096 rows remain, working on a
 * column-major copy of the panel in work (n x 256 bytes), the pivot the
 * first row of largest |re| + |im|), row swaps and the block-row
 * triangular solve per panel, then a tiled FP64 GEMM update of the trailing
 * matrix.  ipiv (n ints): row swapped with row j at step j; perm (n ints):
 * row i of P A is row perm[i] of A; status_dev (zeroed by the caller)
 * receives 1 + the first column whose pivot is zero or non-finite.
 * Replaces jax.scipy.linalg.lu_factor in Matrix.lu_factor
 * (la/matrix.py:137-167), TPMatricesDenseLUFactors (la/tpmatrix.py:1016-1086)
 * and TensorMatrix.solve (la/tensormatrix.py:76-80). */
int jfx_dense_lu_factor(int cplx, void* A, int n, int* ipiv, int* perm, int* status_dev, void* work,
                        hipStream_t stream);
/* X (n x nrhs, row-major) = A^-1 B from the factors above: X = P B, then
 * forward (unit L) and backward (U) substitution.  Real factors with
 * nrhs <= 16: one launch per direction, one workgroup per 64-row block that
 * applies the earlier blocks' results as they are published (device flags,
 * bounded waits; a block that gives up waiting is written as NaN) and solves
 * its diagonal block with reciprocal pivots (3..16 right-hand sides: the
 * block updates on FP64 MFMA).  Otherwise one launch per
 * 32-row block (16 complex) that solves the block in registers and updates
 * the remaining rows.  work holds n x nrhs values of the intermediate
 * L^-1 P B followed by 2 ceil(n/64) ints (the block flags).  B, X and work
 * must not overlap.  Replaces jax.scipy.linalg.lu_solve in Matrix.lu_solve
 * (la/matrix.py:169-214). */
int jfx_dense_lu_solve(int cplx, const void* LU, int n, const int* perm, const void* B, void* X, long long nrhs,
                       void* work, hipStream_t stream);

/* nsys independent band matrices, band (nsys, n, p+q+1) "by row"; status_dev
 * receives 0 or 1 + the row of a zero / non-finite pivot in a failing system.
 * Replaces the per-wavenumber DIA LU of TPMatricesWavenumberSolver
 * (la/tpmatrix.py:735-775, vmapped _lu_banded_no_pivot_kernel). */
int jfx_band_lu_factor_batched(const double* band, int nsys, int n, int p, int q, double* lu, int* status_dev,
                               hipStream_t stream);
/* In-place solve along the middle axis of X (P, n, Q) where line (pp, qq)
 * uses factor (pp * (Q / qdiv) + qq / qdiv) mod nsys (P * Q / qdiv a multiple
 * of nsys, so leading batch dimensions repeat the systems; qdiv = 2 for the
 * re/im pairs of complex data).  Replaces
 * TPMatricesWavenumberSolver.solve (la/tpmatrix.py:590-690, 800-1011). */
int jfx_band_lu_solve_batched(const double* lu, int nsys, int n, int p, int q, double* X, long long P, long long Q,
                              int qdiv, hipStream_t stream);
/* As jfx_band_lu_solve_batched with the kernel chosen by mode, as in
 * jfx_band_lu_solve_mode (0 automatic, 1 one lane per line, 2 wavefronts per
 * line with each line's own factor; jfx_band_lu_solve_batched is mode 0).
 * The same per-wavenumber solve (la/tpmatrix.py:590-690, 800-1011) for few
 * lines. */
int jfx_band_lu_solve_batched_mode(const double* lu, int nsys, int n, int p, int q, double* X, long long P,
                                   long long Q, int qdiv, int mode, hipStream_t stream);

/* One large band matrix (the Kronecker matrix of banded factors), same band
 * storage: LU without pivoting by one 1024-thread workgroup (each
 * elimination step updates its p x q window in parallel; p, q <= 1024);
 * status_dev receives 0 or 1 + the row of a zero / non-finite pivot.
 * Replaces DiaMatrix.lu_factor (la/diamatrix.py:394-504) as reached from
 * TPMatrices.solve(kron_method="banded" | "rcm") (la/tpmatrix.py:429-509,
 * la/diamatrix.py:592-770). */
int jfx_band_lu_factor_wide(const double* band, int n, int p, int q, double* lu, int* status_dev,
                            hipStream_t stream);
/* In-place solves of nrhs right-hand sides X (nrhs, n), rows contiguous, one
 * wave each (n <= 20480, the vector lives in LDS).  Replaces LUFactors.solve
 * (la/diamatrix.py:1880-2195) on the assembled Kronecker matrix. */
int jfx_band_lu_solve_wide(const double* lu, int n, int p, int q, double* X, long long nrhs, hipStream_t stream);

/* out (r m x r m) = sum_t s[t] kron(X_t, Y_t), X (nt, r, r), Y (nt, m, m),
 * s (nt) on the device: the Kronecker assembly of tpmats_to_kron
 * (la/tpmatrix.py:1357-1402). */
int jfx_kron_sum(const double* Xs, const double* Ys, const double* s, int nt, int r, int m, double* out,
                 hipStream_t stream);

/* Y[b, i] = X[b, perm[i]] for X, Y (nrhs, n), perm (n) int64: the RCM
 * reordering of right-hand sides (la/diamatrix.py:687-700). */
int jfx_gather_rows(const double* X, const long long* perm, int n, long long nrhs, double* Y, hipStream_t stream);

/* Complex (re/im interleaved) versions of the two calls above for
 * per-wavenumber systems with complex weights (odd derivative orders on a
 * Fourier axis): band (nsys, n, p+q+1) complex; X (P, n, Q) complex, line
 * (pp, qq) using factor (pp * Q + qq) mod nsys (la/tpmatrix.py:590-1011). */
int jfx_band_lu_factor_batched_z(const double* band, int nsys, int n, int p, int q, double* lu, int* status_dev,
                                 hipStream_t stream);
int jfx_band_lu_solve_batched_z(const double* lu, int nsys, int n, int p, int q, double* X, long long P, long long Q,
                                hipStream_t stream);

/* Y = X * D, complex (re/im interleaved), D (size complex values) broadcast
 * over the leading batch: the diagonal solve of TPMatrix / TPMatrices with
 * complex diagonal factors (la/tpmatrix.py:178-213). */
int jfx_cmul_bcast(const double* X, const double* D, double* Y, long long total, long long size,
                   hipStream_t stream);

/* ---- FD eigenvalue division: Y = X / D with D (size) broadcast over the leading batch: the
 * eigenvalue division of TPMatricesLUFactors.solve (la/tpmatrix.py:565-579). */
int jfx_div_bcast_f64(const double* X, const double* D, double* Y, long long total, long long size,
                      hipStream_t stream);
int jfx_div_bcast_f32(const float* X, const float* D, float* Y, long long total, long long size,
                      hipStream_t stream);

/* Y = X * C with C (size) broadcast over the leading batch: the pointwise
 * coefficient product of a non-separable bilinear form evaluated at the
 * quadrature mesh (assemble_multivar, galerkin/inner.py:991-1024, applied
 * matrix-free as TensorMatrix.matvec, la/tensormatrix.py:54-56). */
int jfx_mul_bcast_f64(const double* X, const double* C, double* Y, long long total, long long size,
                      hipStream_t stream);

/* Y = X + alpha * D with D (size) broadcast over the leading batch: adds the
 * padded boundary-lifting coefficients (or their values on a mesh) of an
 * inhomogeneous Dirichlet space, DirectSum.to_orthogonal / from_orthogonal
 * (galerkin/composite.py:588-602) and DirectSumTPS.to_orthogonal /
 * from_orthogonal (galerkin/tensorproductspace.py:806-851). */
int jfx_axpy_bcast_f64(const double* X, const double* D, double alpha, double* Y, long long total,
                       long long size, hipStream_t stream);

/* Complex result of a complex operator A = Ar + i Ai applied along an axis:
 * Z (P, 2M, Qr) is the real contraction of the stacked operator [Ar; Ai]
 * with the real view of X (Qr = 2Q, re/im interleaved, for complex X; Qr = Q
 * for real X); Y (P, M, Q) complex (re/im interleaved) = Z_top + i Z_bottom.
 * The Fourier transforms' complex exponential series, evaluated and projected
 * by contraction (Fourier.backward / forward / scalar_product,
 * galerkin/Fourier.py:127-180). */
int jfx_cplx_combine(const double* Z, double* Y, long long P, long long M, long long Q, int real_input,
                     hipStream_t stream);

/* y (+)= sum_j coeffs[j] * xs[j], nin <= 8; xs and coeffs are HOST arrays of
 * device pointers / scalars.  dtype 0 = f64, 1 = f32.  The stage
 * combinations of IMEXRungeKutta.stage (integrators/imex_rk.py:55-94). */
int jfx_lincomb(int dtype, int nin, const void* const* xs, const double* coeffs, void* y, long long n,
                int accumulate, hipStream_t stream);

/* out = sum_t coeffs[t] * As[t] * Bs[t] pointwise (nterms <= 4; As, Bs,
 * coeffs are HOST arrays of device pointers / scalars; dtype 0 f64, 1 f32).
 * The quadratic nonlinear terms evaluated on the padded mesh by the IMEX
 * integrators' nonlinear evaluator (integrators/nonlinear.py:89-133). */
int jfx_sum_products(int dtype, int nterms, const void* const* As, const void* const* Bs,
                     const double* coeffs, void* out, long long n, hipStream_t stream);

/* direction 0: f64 -> f32 (round to nearest), 1: f32 -> f64.  The dtype
 * boundary of the mixed-precision transforms (no reference counterpart:
 * the reference runs in one dtype). */
int jfx_cast(int direction, const void* x, void* y, long long n, hipStream_t stream);

/* ---- ragged batch of 2D Legendre-Dirichlet Helmholtz solves
 * (alpha M (x) M - K (x) M - M (x) K) U = F in one persistent launch.
 * items: device array of nitems work items (jfx_helm_item_size() bytes each:
 *   {long long foff, voff; double alpha; int n, p, q, a, b, AP, loff, pad}),
 *   one per (solve, row parity p) (q and b are ignored: the workgroup runs
 *   both column parities in turn).  The items form nqueues (1..16) queues:
 *   queue g is items[qbounds[g] .. qbounds[g+1]) (qbounds: nqueues + 1 device
 *   ints, qbounds[0] = 0, qbounds[nqueues] = nitems), claimed in order through
 *   the device counters[g] (reset by the call itself).  Workgroup w starts on
 *   queue w mod nqueues and moves to the following queues once it is drained;
 *   with one queue per XCD, items that share tables are best kept in one queue.
 * Vd/VTd: FP64 eigenvector tables V_p / V_p^T (AP x AP, zero padded) of the
 *   parity blocks of K_p V = M_p V diag(lam); Vs/VTs: the same as 3 bf16
 *   float-split slices each ([3][AP][AP] per table at 3*voff) in mode 1, or
 *   as the fp16 high and low terms of V_p * 2^texp ([2][AP][AP] at 2*voff,
 *   texp in the item, max |V_p| 2^texp < 2^14) in mode 2.  lam at loff.
 * mode 0: FP64 MFMA (fp64 I/O); mode 1: 3-term BF16 split MFMA with FP32
 *   accumulation/storage and FP64 tridiagonal recurrences whose sweep values
 *   and multipliers are stored in FP32 (io_f32 selects float or double F/U).
 *   nwaves 4 or 8 (mode 1: 4 = two 4-wave workgroups per CU, one column
 *   parity at a time; 8 = one 8-wave workgroup per CU, both column parities
 *   at once with whole-row F reads and U writes); scratch holds
 *   nblocks * jfx_helm_scratch_doubles_per_block() doubles (mode 0 only).
 * mode 2: as mode 1 with a 2-term fp16 split (three MFMA products instead of
 *   six): each operand block is scaled by a power of two from its largest
 *   entry before the split and the products are scaled back exactly; nwaves 4
 *   only. */
 * Replaces TPMatrices.solve -> tpmats_lu_factor / TPMatricesLUFactors.solve
 * (la/tpmatrix.py:429-587, 1089-1233) applied once per solve, and the banded
 * DiaMatrix solve along an axis (la/diamatrix.py:592-770). */
int jfx_helm_item_size(void);
int jfx_helm_scratch_doubles_per_block(void);
int jfx_helm_ragged(int mode, int io_f32, int nwaves, const void* items, int nitems, int* counters,
                    const int* qbounds, int nqueues, const double* Vd, const double* VTd, const unsigned short* Vs,
                    const unsigned short* VTs, const double* lam, const void* F, void* U,
                    double* scratch, int nblocks, hipStream_t stream);

/* Allocation-free form of jfx_helm_ragged in mode 1 (3-term bf16 split,
 * nwaves 4 or 8).  No counterpart in the reference: it serves callers that cannot
 * allocate device memory (code under HIP graph capture, serving loops).  The
 * queue counters live in a static device array of
 * jfx_helm_noalloc_max_queues() (= 16) ints per GPU, cleared by the launch
 * itself; nqueues above that limit is rejected with -2.  No hipMalloc/hipFree
 * is made.  Overlapping calls on one GPU share the counters and race. */
int jfx_helm_noalloc_max_queues(void);
int jfx_helm_ragged_noalloc(int io_f32, int nwaves, const void* items, int nitems, const int* qbounds, int nqueues,
                            const unsigned short* Vs, const unsigned short* VTs, const double* lam, const void* F,
                            void* U, int nblocks, hipStream_t stream);

/* ---- learned preconditioner: no reference counterpart; the
 * nearest analogue is the Linear+activation MLP of pinns/module.py:335-380.
 * C (M x N) = alpha op(A) op(B) + beta C (+ bias[N]) (tanh if act = 1);
 * op(A) = A (M x K, lda) if ta = 0, else A^T of a stored (K x M, lda);
 * op(B) = B (K x N, ldb) if tb = 0, else B^T of a stored (N x K, ldb).
 * Operands are rounded to bf16, products accumulate in FP32 on MFMA. */
int jfx_gemm_bf16(int ta, int tb, int M, int N, int K, float alpha, const float* A, int lda, const float* B,
                  int ldb, float beta, float* C, int ldc, const float* bias, int act, hipStream_t stream);
/* dZ = dY (1 - H^2) */
int jfx_tanh_backward(const float* dY, const float* H, float* dZ, long long n, hipStream_t stream);
/* out[j] = sum_i X[i, j] */
int jfx_colsum(const float* X, float* out, int rows, int cols, hipStream_t stream);
/* Adam step `step` (>= 1) with bias correction. */
int jfx_adam(float* p, const float* g, float* m, float* v, long long n, float lr, float b1, float b2, float eps,
             int step, hipStream_t stream);

/* ---- batched PCG (one SPD system per batch entry, FP64; the operator and
 * preconditioner applications are contractions above).  The reference has
 * no PCG; the operator is TensorMatrix's matvec (la/tensormatrix.py:54-56).
 * out[b] = <x_b, y_b> */
int jfx_batched_dot(const double* x, const double* y, double* out, int batch, long long n, hipStream_t stream);
/* x_b += (num_b/den_b) p_b, r_b -= (num_b/den_b) q_b for active b */
int jfx_pcg_update_xr(double* x, double* r, const double* p, const double* q, const double* num,
                      const double* den, const int* active, int batch, long long n, hipStream_t stream);
/* p_b = z_b + (num_b/den_b) p_b for active b */
int jfx_pcg_update_p(double* p, const double* z, const double* num, const double* den, const int* active,
                     int batch, long long n, hipStream_t stream);
/* active_b &= rr_b > tol2 bb_b, counting iters_b for systems active on entry when count != 0;
 * a non-finite rr_b stops system b with iters_b = -1 (breakdown) */
int jfx_pcg_check(const double* rr, const double* bb, double tol2, int* active, int* iters, int batch, int count,
                  hipStream_t stream);
/* y_b = s_b x_b */
int jfx_scale_batch(const double* x, const double* s, double* y, int batch, long long n, hipStream_t stream);
/* y = exp(scale x): the coefficient field k = exp(g) at the quadrature mesh */
int jfx_exp(const double* x, double* y, double scale, long long n, hipStream_t stream);

#ifdef __cplusplus
}
#endif
#endif /* JFX_H */
