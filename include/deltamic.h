/*
 * deltamic_amd C-ABI: the narrow-band analytic Fourier transform of triangle
 * meshes ("meshFT") and its adjoint, as hand-written HIP kernels for gfx950.
 *
 * Every entry point takes plain device pointers and sizes, enqueues its
 * kernels on `stream` (no device-wide synchronisation, no allocation: scratch
 * memory is a caller-provided `workspace` whose size the matching *_workspace
 * function returns) and returns DMIC_OK or a negative error code whose text
 * dmic_last_error() returns.  Results are deterministic: no floating-point
 * atomics; every reduction has a fixed order.
 *
 * Layouts (all row-major, device memory):
 *   verts       float [nv][3]   vertex positions already shifted to the box
 *                                minimum (render.py:148 does Verts - box_size[:,0])
 *   faces       int64 or int32 [nf][3] (faces_is_int64 selects)
 *   faces_coeff float [nf]
 *   xi_list     float [n][3]    compacted band frequencies
 *   ftmesh_list float [n][2]    complex64 band values (re, im)
 *
 * Face indices outside [0, nv) do not fault: the vertex reads as the origin
 * and DMIC_ERR_INDEX is reported by dmic_check_index_flag().
 */
#ifndef DELTAMIC_AMD_H
#define DELTAMIC_AMD_H

#include <stddef.h>
#include <stdint.h>
#include <hip/hip_runtime_api.h>

#ifdef __cplusplus
extern "C" {
#endif

#define DMIC_OK 0
#define DMIC_ERR_ARG (-1)
#define DMIC_ERR_WORKSPACE (-2)
#define DMIC_ERR_HIP (-3)
#define DMIC_ERR_INDEX (-4)

/* Library version (major*10000 + minor*100 + patch). */
int dmic_version(void);

/* Text of the last error reported on the calling thread ("" if none). */
const char* dmic_last_error(void);

/* ------------------------------------------------------------------ meshFT */

/* Workspace bytes dmic_meshft_forward needs for these sizes. */
size_t dmic_meshft_forward_workspace(int64_t nv, int64_t nf, int64_t n);

/*
 * ftmesh_list[m] = sum_t w_t F(xi_m . v_a, xi_m . v_b, xi_m . v_c),
 *   w_t = 2 c_t A_t / sum_s c_s A_s,  F = 2nd divided difference of -exp(-ix).
 * Replaces forward_fourier3d_cpp -> fourier3d_forward -> fourier3d_cuda_forward
 * (deltamic/cpp/fourier3d_cuda.cpp:27-42, fourier3d_cuda_kernel.cu:421-462).
 * n == 0 or nf == 0 is valid (zero output / nothing to do).
 */
int dmic_meshft_forward(const float* verts, int64_t nv, const void* faces, int faces_is_int64,
                        int64_t nf, const float* faces_coeff, const float* xi_list, int64_t n,
                        float* ftmesh_list, void* workspace, size_t workspace_bytes,
                        hipStream_t stream);

/*
 * float64 boundary of dmic_meshft_forward / dmic_meshft_backward (the
 * reference's AT_DISPATCH_FLOATING_TYPES double case,
 * fourier3d_cuda_kernel.cu:447,508): vertices, faces_coeff, xi_list and the
 * upstream gradient are double, ftmesh_list is complex128 (double pairs) and
 * grad_verts double.  Inside, the same kernels run: vertex phases in FP64
 * from the FP64 inputs, edge vectors and the face formula in FP32, FP64
 * accumulation; the result carries the float32 path's accuracy (about 4e-7
 * forward / 4e-6 adjoint rel-L2 against an FP64 evaluation).  Same
 * workspaces as the float32 entry points.
 */
int dmic_meshft_forward_f64(const double* verts, int64_t nv, const void* faces,
                            int faces_is_int64, int64_t nf, const double* faces_coeff,
                            const double* xi_list, int64_t n, double* ftmesh_list,
                            void* workspace, size_t workspace_bytes, hipStream_t stream);
int dmic_meshft_backward_f64(const double* verts, int64_t nv, const void* faces,
                             int faces_is_int64, int64_t nf, const double* faces_coeff,
                             const double* grad_output_list, const double* ftmesh_list,
                             const double* xi_list, int64_t n, double* grad_verts,
                             void* workspace, size_t workspace_bytes, hipStream_t stream);

/*
 * dmic_meshft_forward without a workspace and without any allocation.  No
 * counterpart in the reference: it is for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  Its scratch memory (packed face records
 * and area sums, 72 B per face) is a static 20 MiB device array that each
 * GPU the library runs on holds until the process ends.  Limit:
 * nf <= 262144 faces (DMIC_ERR_ARG beyond); any band size n.  Every call on one GPU
 * shares that GPU's array, so calls that overlap on one GPU (other streams,
 * other threads) race and must be serialised by the caller.  A face index
 * outside [0, nv) reads as the origin and is not reported (there is no
 * workspace for dmic_check_index_flag); validate indices beforehand.
 */
int dmic_meshft_forward_noalloc(const float* verts, int64_t nv, const void* faces,
                                int faces_is_int64, int64_t nf, const float* faces_coeff,
                                const float* xi_list, int64_t n, float* ftmesh_list,
                                hipStream_t stream);

/*
 * Workspace bytes dmic_meshft_backward (and dmic_meshft_backward_f64) needs
 * for these sizes on the current device: the adjoint's frequency chunks are
 * sized from the device's resident workgroups (80 B per face and chunk, plus
 * 8 B per frequency), so query with the device that will run the call
 * current.
 */
size_t dmic_meshft_backward_workspace(int64_t nv, int64_t nf, int64_t n);

/*
 * grad_verts[v] = d/dv sum_m Re(conj(g_m) ftmesh_m), g = grad_output_list
 * (torch's complex gradient convention, fourier3d_cuda_kernel.cu:392-394).
 * Per-face partial sums over frequency tiles, then a fixed-order CSR
 * face->vertex reduction (any vertex valence).  ftmesh_list is accepted for
 * signature compatibility; the area-normalisation term is rebuilt from the
 * per-face sums.  Replaces backward_fourier3d_cpp -> fourier3d_backward ->
 * fourier3d_cuda_backward (fourier3d_cuda.cpp:44-65, fourier3d_cuda_kernel.cu:466-539).
 */
int dmic_meshft_backward(const float* verts, int64_t nv, const void* faces, int faces_is_int64,
                         int64_t nf, const float* faces_coeff, const float* grad_output_list,
                         const float* ftmesh_list, const float* xi_list, int64_t n,
                         float* grad_verts, void* workspace, size_t workspace_bytes,
                         hipStream_t stream);

/*
 * Autofluorescence variant of the forward: every face weighs
 *   W_t(xi) = w_t + i (xi . nor2_t) / |xi|^2,
 *   nor2_t = regint / regvol * ((b - a) x (c - a)),  regvol = -sum_t a . (b x c) / 6,
 * so ftmesh(xi) = sum_t W_t F_t(xi) is the surface signal plus a uniform
 * interior signal of strength regint (divergence theorem).  At xi = 0 the
 * interior term is its limit -regint (the reference divides by zero there).
 * Same arguments as dmic_meshft_forward plus regint; the workspace is
 * dmic_meshft_forward_workspace's.  Replaces the dense torch function
 * fourier_from_mesh_narrowband_autofluorescence (deltamic/render.py:228-318;
 * it has no Faces_coeff: pass ones for its weights 2 A_t / sum A).
 */
int dmic_meshft_af_forward(const float* verts, int64_t nv, const void* faces, int faces_is_int64,
                           int64_t nf, const float* faces_coeff, float regint,
                           const float* xi_list, int64_t n, float* ftmesh_list, void* workspace,
                           size_t workspace_bytes, hipStream_t stream);

/* Workspace bytes dmic_meshft_af_backward needs for these sizes. */
size_t dmic_meshft_af_backward_workspace(int64_t nv, int64_t nf, int64_t n);

/*
 * grad_verts = d/dV sum_m Re(conj(g_m) ftmesh_m) of the autofluorescence
 * variant, including the dependence of nor2_t and regvol on the vertices
 * (the reference differentiates its dense expression with autograd).
 */
int dmic_meshft_af_backward(const float* verts, int64_t nv, const void* faces, int faces_is_int64,
                            int64_t nf, const float* faces_coeff, float regint,
                            const float* grad_output_list, const float* xi_list, int64_t n,
                            float* grad_verts, void* workspace, size_t workspace_bytes,
                            hipStream_t stream);

/*
 * Reads (synchronously, on `stream`) the out-of-range face index flag that
 * the last dmic_meshft_forward / dmic_meshft_backward call on `workspace`
 * set (each call clears it first).  Returns DMIC_OK or DMIC_ERR_INDEX.  No
 * counterpart in the reference, which does not check indices.
 */
int dmic_check_index_flag(void* workspace, hipStream_t stream);

/*
 * Per-kernel timing for benchmarks (no counterpart in the reference): when
 * enabled, hipEvents are recorded on the launch stream around the main
 * forward kernel and the main adjoint kernel of every call.  Enabling also
 * resets the totals.  dmic_profile_read waits for the last recorded launch
 * and returns the summed milliseconds and launch counts.  Not thread-safe.
 */
int dmic_profile_enable(int enable);
int dmic_profile_read(double* fwd_ms, long* fwd_count, double* bwd_ms, long* bwd_count);

/* -------------------------------------------------------- narrow band */

/* Workspace bytes for dmic_band_count / dmic_band_fill on an n0 x n1 x n2 box. */
size_t dmic_band_workspace(int64_t n0, int64_t n1, int64_t n2);

/*
 * Counts the band voxels filter[i] > thresh (no abs, as cuda_class.py:21) and
 * writes the count to *count_dev (device int64).  Together with
 * dmic_band_fill this replaces the boolean-mask gather of cuda_class.py:20-23.
 */
int dmic_band_count(const float* filter, int64_t n0, int64_t n1, int64_t n2, float thresh,
                    int64_t* count_dev, void* workspace, size_t workspace_bytes,
                    hipStream_t stream);

/*
 * After dmic_band_count on the same workspace: writes the band in row-major
 * voxel order: idx[j] = linear voxel index, xi_list[j] = (xi0[i0], xi1[i1],
 * xi2[i2]) (the meshgrid 'ij' gather of cuda_class.py:20-23).
 */
int dmic_band_fill(const float* filter, int64_t n0, int64_t n1, int64_t n2, float thresh,
                   const float* xi0, const float* xi1, const float* xi2, int32_t* idx,
                   float* xi_list, void* workspace, size_t workspace_bytes, hipStream_t stream);

/* box[:] = 0 then box[idx[j]] = vals[j]: the scatter ftmesh[mask] = ftmesh_list
 * of cuda_class.py:27-28.  vals and box are complex64 as float pairs. */
int dmic_band_scatter(const float* vals, const int32_t* idx, int64_t n, float* box, int64_t nbox,
                      hipStream_t stream);

/* vals[j] = box[idx[j]]: the gather grad_output[mask] of cuda_class.py:36. */
int dmic_band_gather(const float* box, const int32_t* idx, int64_t n, float* vals,
                     hipStream_t stream);

/*
 * Hermitian halving (no counterpart in the reference, which evaluates every
 * band frequency).  On the reference's grid the axes are exactly
 * antisymmetric (xi_d[k] = -xi_d[n_d-1-k], render.py:163-171); if the band
 * mask is centrosymmetric as well, band entry n-1-j is the mirror -xi of
 * entry j and ftmesh(-xi) = conj(ftmesh(xi)) for a real mesh.
 *
 * dmic_band_symmetry ORs into *flag_dev (device int, caller zeroes it):
 * 1 if the mask is not centrosymmetric, 2 if an axis is not antisymmetric.
 * dmic_band_mirror_fill: vals[n-1-j] = conj(vals[j]) for j < n/2, after the
 * forward has written the first (n+1)/2 entries.
 * dmic_band_mirror_fold: g_half[j] = g[j] + conj(g[n-1-j]) for j < n/2 (the
 * middle entry of an odd band is copied): the upstream gradient of the
 * half band, so the adjoint runs on (n+1)/2 frequencies.
 */
int dmic_band_symmetry(const float* filter, int64_t n0, int64_t n1, int64_t n2, float thresh,
                       const float* xi0, const float* xi1, const float* xi2, int* flag_dev,
                       hipStream_t stream);
int dmic_band_mirror_fill(float* vals, int64_t n, hipStream_t stream);
int dmic_band_mirror_fold(const float* g, int64_t n, float* g_half, hipStream_t stream);

/*
 * out[i0][i1][i2] = exp(-xi^T Sigma xi / 2), xi = (xi0[i0], xi1[i1], xi2[i2]);
 * sigma is a HOST array of 9 floats (row-major 3x3).  Replaces
 * generate_gaussian_psf (deltamic/psf.py:15-20).
 */
int dmic_gaussian_otf(const float* sigma_host, const float* xi0, const float* xi1,
                      const float* xi2, int64_t n0, int64_t n1, int64_t n2, float* out,
                      hipStream_t stream);

/* ------------------------------------------------- learnable PSFs
 *
 * Replace the forward of Gaussian_psf / Gaussian_psf_logsigma
 * (deltamic/psf.py:551-625) and Anisotropic_gaussian_psf (psf.py:628-665)
 * and the autograd backward torch derives for them.  Parameters live in
 * DEVICE memory (they are learnable tensors).  Axes x0[n0], x1[n1], x2[n2];
 * the PSF is [n0][n1][n2].
 */

/* psf = exp(-x^T S x / 2); S: 9 floats, row-major 3x3 (psf.py:572-576). */
int dmic_psf_gaussian(const float* S, const float* x0, const float* x1, const float* x2,
                      int64_t n0, int64_t n1, int64_t n2, float* out, hipStream_t stream);

/* Workspace bytes for the two gradient entry points below. */
size_t dmic_psf_grad_workspace(int64_t n0, int64_t n1, int64_t n2);

/* grad_S[9] (device) = dL/dS given grad_out = dL/dpsf and the psf of the
 * forward.  Fixed-order FP64 reduction: bitwise reproducible. */
int dmic_psf_gaussian_grad(const float* grad_out, const float* psf, const float* x0,
                           const float* x1, const float* x2, int64_t n0, int64_t n1, int64_t n2,
                           float* grad_S, void* workspace, size_t workspace_bytes,
                           hipStream_t stream);

/* psf = exp(-Q / (2D)) / D, Q = (x0, x1) S2 (x0, x1)^T, D = eps + (|x2| sz)^2
 * (psf.py:655-660); prm = {S2_00, S2_01, S2_10, S2_11, sz, eps}. */
int dmic_psf_aniso(const float* prm, const float* x0, const float* x1, const float* x2,
                   int64_t n0, int64_t n1, int64_t n2, float* out, hipStream_t stream);

/* grad_prm[6] (device) = dL/dprm, same layout as prm. */
int dmic_psf_aniso_grad(const float* prm, const float* grad_out, const float* psf,
                        const float* x0, const float* x1, const float* x2, int64_t n0, int64_t n1,
                        int64_t n2, float* grad_prm, void* workspace, size_t workspace_bytes,
                        hipStream_t stream);

/* ------------------------------------------------- render chain
 *
 * The modulus and normalisation after the inverse FFT of
 * render_image_from_ftmesh (deltamic/render.py:177-180) and normalize_tensor
 * (render.py:5-6), forward and backward, as HBM-bound passes.  z is a
 * complex64 box (float pairs), t / img / out float; minmax is a device
 * float[2] (min, max); NaN propagates through min and max.  The workspace
 * (dmic_render_workspace) holds fixed-order reduction partials.
 */
size_t dmic_render_workspace(int64_t n);
/* img = |z| (hypot) and minmax = (min img, max img) in the same pass. */
int dmic_abs_minmax(const float* z, int64_t n, float* img, float* minmax, void* workspace,
                    size_t workspace_bytes, hipStream_t stream);
/* minmax = (min t, max t). */
int dmic_minmax(const float* t, int64_t n, float* minmax, void* workspace, size_t workspace_bytes,
                hipStream_t stream);
/* out = ((t - min) / (max - min) + c) / (1 + c) in FP32, torch's operation order. */
int dmic_normalize(const float* t, int64_t n, const float* minmax, float c, float* out,
                   hipStream_t stream);
/* grad_t = dL/dt of normalize given g = dL/dout, including the paths through
 * min and max (each shared evenly by the entries that attain it, as torch). */
int dmic_normalize_backward(const float* g, const float* t, int64_t n, const float* minmax,
                            float c, float* grad_t, void* workspace, size_t workspace_bytes,
                            hipStream_t stream);
/* grad_z = g z / |z| (0 where z = 0): the backward of |z|. */
int dmic_abs_backward(const float* g, const float* z, int64_t n, float* grad_z,
                      hipStream_t stream);

/* ------------------------------------------------- large-steps preconditioner
 *
 * x = (I + lambda L)^-power g for the vertex gradient g (float [nv][3]),
 * L = D - A the combinatorial Laplacian of the mesh's vertex graph given as
 * CSR adjacency (row_ptr[nv + 1], col[row_ptr[nv]]: each vertex's
 * neighbours, no self loops; device int32).  Replaces the reference fit
 * loop's sparse Cholesky solve of (I + lambda L)^2 (Gradient_Descent.ipynb
 * cell 2: compute_matrix, CholeskySolver(M @ M), solver.solve) with `power`
 * Jacobi-preconditioned conjugate-gradient solves of M = I + lambda L to a
 * relative residual `tol` (at most max_iter iterations each) in one
 * workgroup; FP64 vectors, fixed-order reductions (bitwise reproducible).
 * iters_out (optional, device int) receives the last solve's iterations.
 */
size_t dmic_laplacian_solve_workspace(int64_t nv);
int dmic_laplacian_solve(const int* row_ptr, const int* col, int64_t nv, float lambda, int power,
                         const float* g, float* x, float tol, int max_iter, int* iters_out,
                         void* workspace, size_t workspace_bytes, hipStream_t stream);

/* ------------------------------------------------- proposal network (MLP)
 *
 * No counterpart in the reference: the learned per-vertex displacement
 * proposal (15 -> 128 -> 128 -> 128 -> 3, ReLU) of the fitting loop.  Weights
 * are bf16 (packed row-major [in][out]: W1 15x128, W2 128x128, W3 128x128,
 * W4 128x3), biases fp32 (b1, b2, b3: 128, b4: 3); products accumulate in
 * FP32 on the matrix cores (v_mfma_f32_16x16x16_bf16).
 */

/* C[M,N] = A(M,K) B(K,N) on MFMA with bf16 inputs and FP32 accumulation;
 * A(i,k) = A[i*sa0 + k*sa1], B(k,j) = B[k*sb0 + j*sb1] (fp32 or bf16 storage).
 * epilogue: 0 none, 1 + bias[j], 2 relu(+ bias[j]), 3 zero where mask[i*ldm+j] <= 0. */
int dmic_gemm_bf16(const void* A, int a_is_bf16, int64_t sa0, int64_t sa1, const void* B,
                   int b_is_bf16, int64_t sb0, int64_t sb1, float* C, int64_t ldc, int64_t M,
                   int64_t N, int64_t K, const float* bias, const float* mask, int64_t ldm,
                   int epilogue, hipStream_t stream);

int64_t dmic_mlp_weight_count(void);
int64_t dmic_mlp_bias_count(void);

/* y[rows][3] = MLP(x[rows][15]).  acts (optional, [3][rows][128]) keeps the
 * hidden activations for dmic_mlp_backward; without it scratch must hold
 * 2*rows*128 floats. */
int dmic_mlp_forward(const uint16_t* w_bf16, const float* bias, const float* x, int64_t rows,
                     float* y, float* acts, float* scratch, hipStream_t stream);

/* Gradients of sum(dy . y) w.r.t. the weights (grad_w, fp32) and biases
 * (grad_b); scratch holds 2*rows*128 floats.  Deterministic. */
int dmic_mlp_backward(const uint16_t* w_bf16, const float* x, const float* acts, const float* dy,
                      int64_t rows, float* grad_w, float* grad_b, float* scratch,
                      hipStream_t stream);

/* Bias-corrected Adam on fp32 master parameters; w_bf16 (optional) receives
 * the updated bf16 copy. */
int dmic_adam_step(float* w, uint16_t* w_bf16, const float* g, float* m, float* v, int64_t n,
                   float lr, float beta1, float beta2, float eps, int64_t step,
                   hipStream_t stream);

/* y = bf16(x), round to nearest even. */
int dmic_to_bf16(const float* x, int64_t n, uint16_t* y, hipStream_t stream);

#ifdef __cplusplus
}
#endif

#endif /* DELTAMIC_AMD_H */
