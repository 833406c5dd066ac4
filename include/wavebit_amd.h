/* wavebit_amd -- C ABI of the MI355X RCWA engine (gfx950, complex FP64).
 *
 * Every buffer is caller-allocated device memory (for example a torch tensor's data_ptr()),
 * complex128 as interleaved {re, im} doubles, matrices row-major. The library never frees
 * caller memory and allocates no device memory of its own: each entry point that needs
 * scratch takes a workspace pointer and byte count, sized by its *_workspace_size query.
 * No C++ exception crosses this boundary; every call returns a wb_status and leaves a
 * message for wb_last_error(). `stream` is a hipStream_t (NULL = the null stream). Work is
 * enqueued asynchronously; outputs are valid once the stream has reached the call.
 *
 * A context belongs to one host thread and one device; share it across threads only under an
 * external lock.
 *
 * The reference (fieldalgo/wavebit) is pure Python; the functions below replace the internal
 * seams its raster RCWA path calls (paths relative to the reference's src/wavebit/):
 *   wb_assemble_li   rcwa/_engine/factorization.py:915-940  to_conv_mat_raster_sampled
 *                    (li_conv_matrices, :658-671), one call for a whole layer stack;
 *   wb_assemble_laurent
 *                    rcwa/_engine/factorization.py:1016-1034 to_conv_mat_raster_laurent
 *                    (laurent_conv_matrices, :955-1000; _laurent_conv, :944-952), the
 *                    Numerics(factorization='laurent') opt-in, one call for a layer stack;
 *   wb_assemble_nvf  rcwa/_engine/factorization.py:2840-2915 to_conv_mat_raster_nvf
 *                    (nvf_conv_matrices, :2796-2815; normal_vector_field, :1180-1404; the
 *                    closing _nvf_assemble, :2461-2540), Numerics(factorization='nvf') with
 *                    the default normal field ('distance') and projector ('linear');
 *   wb_solve         rcwa/_engine/scattering.py:1386-1456 smm_2d and :704-902 smm_1d as the
 *                    engine calls them (rcwa/_engine/engine.py:3283-3316, 3336-3434), for a
 *                    batch of plane-wave points instead of one wavelength per call
 *                    (rcwa/_lifecycle.py:712-727);
 *   wb_solve_scattering
 *                    the same solve with Outputs(scattering=True): the stack's ScatteringMatrix
 *                    blocks as rcwa/_engine/scattering.py:607-656 (_scattering_matrix,
 *                    _coupled_ports_1d) and rcwa/_engine/engine.py:2995-3009 (cell-centre signs)
 *                    publish them, read by rcwa/results.py:136-141 (Result.scattering);
 *   wb_zgeev_batched _common/backend.py:1767-1812 eig (numpy.linalg.eig / torch.linalg.eig);
 *   wb_zgetrf_batched / wb_zgetrs_batched / wb_zgemm_batched
 *                    the dense inv / solve / matmul calls of rcwa/_engine/scattering.py:116-200
 *                    (_ab_matrices, _s_layer, redheffer_star) and _common/backend.py:506-542
 *                    (linear_solve).
 *   wb_solve_keep / wb_solve_adjoint / wb_assemble_li_adjoint / wb_assemble_laurent_adjoint /
 *   wb_assemble_nvf_adjoint
 *                    the reverse pass PyTorch autograd takes through the same path
 *                    (loss.backward() after Prepared.run, rcwa/_lifecycle.py:785-786), with
 *                    the eigensolve's generator rule of rcwa/_engine/modal_gradient.py:161-251
 *                    and the divided-difference exponential of :446-470 written by hand; under
 *                    NVF through the builder of factorization.py:2796-2915 with the normal
 *                    field held fixed (tests/test_frozen_geometry.py:644);
 *   wb_zgemm_batched_ops / wb_zgetrs_batched_ops
 *                    the conjugate-transposed products and solves of that reverse pass
 *                    (torch matmul / linalg.solve backward).
 */
#ifndef WAVEBIT_AMD_H
#define WAVEBIT_AMD_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct wb_ctx wb_ctx;

typedef enum {
  WB_OK = 0,
  WB_BADARG = 1,      /* an argument is out of range or inconsistent */
  WB_OOM = 2,         /* the workspace given is smaller than the query returned */
  WB_SINGULAR = 3,    /* a factorization was singular (wb_flags_status of a batch's flags) */
  WB_NONFINITE = 4,   /* reserved for callers that check outputs */
  WB_NOCONVERGE = 5,  /* an eigensolve did not converge (wb_flags_status of a batch's flags) */
  WB_FLAGGED = 6,     /* a pair carries a bit of the caller's notice mask (wb_flags_status) */
  WB_HIP = 7,         /* a HIP runtime call failed; see wb_last_error */
  WB_UNSUPPORTED = 8  /* the combination asked for is not implemented (e.g. a 1D adjoint) */
} wb_status;

/* The status of a batch from the per-pair flags a solve wrote (host copy, count entries):
 * WB_NOCONVERGE if any pair's eigensolve did not converge, else WB_SINGULAR if any factorization
 * was singular, else WB_FLAGGED if any pair carries a bit of notice_mask (e.g. WB_FLAG_GRAZING |
 * WB_FLAG_EP_LOUD), else WB_OK. The solve entries themselves return WB_OK whenever they ran:
 * a batch holds independent points and the reference decides per point (its solves take one
 * wavelength each), so numerical outcomes are reported per (point, layer) pair in flags and a
 * failed point never fails its batch; this helper is the batch-level reading for C callers.
 * Needs no context or device. */
int wb_flags_status(const int* flags, long long count, int notice_mask);

/* wb_solve mode */
enum { WB_MODE_2D = 0, WB_MODE_1D_TE = 1, WB_MODE_1D_TM = 2 };

/* per-point record layout for wb_solve (doubles, WB_PT_STRIDE per point) */
enum {
  WB_PT_WAVELENGTH = 0,   /* vacuum wavelength, same length unit as the periods */
  WB_PT_KX0 = 1,          /* incident in-plane wavevector / k0 (real angles) */
  WB_PT_KY0 = 2,
  WB_PT_PHI = 3,          /* frame azimuth at k_par = 0 (radians) */
  WB_PT_PERIOD_X = 4,
  WB_PT_PERIOD_Y = 5,
  WB_PT_EPS_TOP = 6,      /* superstrate permittivity (re, im): the lit half-space */
  WB_PT_EPS_BOTTOM = 8,   /* substrate permittivity (re, im) */
  WB_PT_A_S = 10,         /* incident Jones pair, unit norm (re, im) each */
  WB_PT_A_P = 12,
  WB_PT_STRIDE = 16
};

/* flag bits written per (point, layer) by wb_solve
 * WB_FLAG_EP: a 2D layer's eigensolve returned a near-defective mode pair (two modes that are
 * each other's nearest eigenvalue with unit eigenvectors at a sine below 1e-2, roots above
 * 1e-9; rebuilt by the reference under exceptional_points='rebuild', rcwa/_engine/
 * solve_2d.py:583-660). WB_FLAG_EP_LOUD: such a pair at a sine below sqrt(2^-53 / 1e-10) =
 * 1.05e-3, where the unrebuilt modes may be off by more than 1e-10 (the reference measures the
 * error as K u / s^2 with K <= 1e-2, :222-229; this bound takes K = 1): a caller holding the
 * reference's tolerance refuses the point. Above that sine the modes as computed are within it.
 * WB_FLAG_GRAZING: a 2D layer has a grazing mode, |w| below 1e-9 of the largest (always rebuilt
 * by the reference, :2463-2485): wb_solve_rebuild solves such points again with the cluster
 * re-derived, and then sets WB_FLAG_REBUILT (rebuilt, pair flags cleared) or WB_FLAG_DEFECTIVE
 * (the cluster is numerically defective: no mode basis exists there; refuse the point). A grazing
 * layer with neither kept its modes (cluster above 8 modes, or its two products disagree). */
enum {
  WB_FLAG_NOCONVERGE = 1, WB_FLAG_SINGULAR = 2, WB_FLAG_EP = 4, WB_FLAG_GRAZING = 8,
  WB_FLAG_REBUILT = 16, WB_FLAG_DEFECTIVE = 32, WB_FLAG_EP_LOUD = 64
};

int wb_version(void);
int wb_ctx_create(int device, wb_ctx** out);
int wb_ctx_destroy(wb_ctx* ctx);
const char* wb_last_error(const wb_ctx* ctx);

/* C[b] = alpha A[b] B[b] + beta C[b]; A m x k, B k x n, C m x n; alpha/beta point to two
 * doubles on the host. stride* are element strides between batch entries (0 = shared). */
int wb_zgemm_batched(wb_ctx* ctx, int m, int n, int k, const double* alpha, const double* A,
                     int lda, long long strideA, const double* B, int ldb, long long strideB,
                     const double* beta, double* C, int ldc, long long strideC, int batch,
                     void* stream);

/* LU with partial pivoting in place; ipiv: batch x n (0-based row swapped with row j);
 * info: batch ints, 0 or j+1 for the first exactly-zero pivot. */
int wb_zgetrf_batched(wb_ctx* ctx, int n, double* A, int lda, long long strideA, int* ipiv,
                      int* info, int batch, void* stream);

/* Solve A X = B in place in B from wb_zgetrf_batched's factors. */
int wb_zgetrs_batched(wb_ctx* ctx, int n, int nrhs, const double* LU, int lda, long long strideA,
                      const int* ipiv, double* B, int ldb, long long strideB, int batch,
                      void* stream);

/* General complex eigendecomposition A W = W diag(w) of `batch` contiguous n x n matrices
 * (A is destroyed). W columns have unit 2-norm with their largest component real.
 * info: batch ints, 0 or i+1 when the QR iteration did not converge at eigenvalue i. */
size_t wb_zgeev_workspace_size(int n, int batch);
int wb_zgeev_batched(wb_ctx* ctx, int n, double* A, double* w, double* W, void* workspace,
                     size_t workspace_bytes, int* info, int batch, void* stream);

/* Li-rule convolution matrices (sampled coefficients, fixed nesting) of L raster layers
 * eps (L, Ny, Nx): epx, epy, epz_i, each (L, N, N), N = (2hu+1)(2hv+1). hv <= 15.
 * info: 2L ints (inversion of the outer epx operator per layer, then of the direct rule). */
size_t wb_assemble_li_workspace_size(int L, int Ny, int Nx, int hu, int hv);
int wb_assemble_li(wb_ctx* ctx, const double* eps, int L, int Ny, int Nx, int hu, int hv,
                   double* epx, double* epy, double* epz_i, void* workspace,
                   size_t workspace_bytes, int* info, void* stream);

/* Laurent-rule convolution matrices (sampled coefficients) of L raster layers eps (L, Ny, Nx):
 * epx = epy = the direct rule [[eps]] (equal to the bit), epz_i = [[eps]]^-1 (the same matrix
 * wb_assemble_li returns in that slot). hu, hv <= 15. info: L ints (inversion per layer). */
size_t wb_assemble_laurent_workspace_size(int L, int Ny, int Nx, int hu, int hv);
int wb_assemble_laurent(wb_ctx* ctx, const double* eps, int L, int Ny, int Nx, int hu, int hv,
                        double* epx, double* epy, double* epz_i, void* workspace,
                        size_t workspace_bytes, int* info, void* stream);

/* Normal-vector-field convolution matrices (sampled coefficients) of L raster layers eps
 * (L, Ny, Nx): with E = [[eps]], A = [[1/eps]]^-1, Delta = E - A and C_ab the point-sample
 * series of n_a n_b (n the unit normal of the periodic distance to the interfaces, on the raster
 * block-repeated to 4h + 1 samples per axis),
 *   epx = E - (Delta C_xx + C_xx Delta)/2, epy = E - (Delta C_yy + C_yy Delta)/2,
 *   epxy = (Delta C_xy + C_xy Delta)/2 (the solve's epxy and epyx slot alike), epz_i = E^-1.
 * lattice: host array (a1x, a1y, a2x, a2y), the cell basis the raster indexes (the normal is
 * taken in Cartesian components through the cell's metric), or NULL for square pixels in an
 * axis-aligned cell. floor = 1: on every layer holding a value that is not a passive dielectric
 * (Re eps <= 0 or Im eps < 0), the singular values of [[1/eps]] below 1 / (2 max|eps|) are
 * raised to it before the closing (factorization.py:1968-2160; one N x N Hermitian Schur
 * decomposition per layer of the call, the passive layers' matrices unchanged); floor = 0 skips
 * that work for a caller that knows every layer is passive. hu, hv <= 15. info: 3 L device ints:
 * the inversions of E, of [[1/eps]], then 1 for a layer that is not a passive dielectric. */
size_t wb_assemble_nvf_workspace_size(int L, int Ny, int Nx, int hu, int hv);
int wb_assemble_nvf(wb_ctx* ctx, const double* eps, int L, int Ny, int Nx, int hu, int hv,
                    const double* lattice, int floor, double* epx, double* epy, double* epz_i,
                    double* epxy, void* workspace, size_t workspace_bytes, int* info, void* stream);

/* Batch solve of P plane-wave points, each through L layers (bottom to top), top incidence.
 *   pts       P x WB_PT_STRIDE doubles (layout above)
 *   thick     P x L layer thicknesses
 *   conv_idx  P x L: which (epx, epy, epz_i) matrix set layer l of point p uses
 *   epx, epy, epz_i   C x N x N complex
 *   epxy, epyx        C x N x N complex, the off-diagonal slots of an NVF layer (2D path;
 *                     the same array twice for wb_assemble_nvf's symmetric projector), or both
 *                     NULL for Li's and Laurent's rules (smm_2d's epxy_conv_all /
 *                     epyx_conv_all, rcwa/_engine/solve_2d.py:2393-2411)
 * Outputs:
 *   eff  P x 3 x 4 x N doubles: incidence variant (as given, pure s, pure p) x (R_s, R_p, T_s,
 *        T_p) per order, orders flattened row-major over (2hv+1, 2hu+1). 1D modes fill
 *        variant 0 only.
 *   amp  P x 3 x 4 x N complex: (r_s, r_p, t_s, t_p) Fresnel amplitudes, centre-referenced
 *   pe   P x 3 power entering the lit face / incident power
 *   flags P x L ints (WB_FLAG_*) or NULL */
size_t wb_solve_workspace_size(int P, int L, int hu, int hv, int mode);
int wb_solve(wb_ctx* ctx, int P, int L, int hu, int hv, int mode, const double* pts,
             const double* thick, const int* conv_idx, const double* epx, const double* epy,
             const double* epz_i, const double* epxy, const double* epyx, double* eff,
             double* amp, double* pe, int* flags,
             void* workspace, size_t workspace_bytes, void* stream);

/* wb_solve without a workspace argument, for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc). No counterpart
 * in the reference (which allocates through NumPy/PyTorch on every call). The solve runs on a
 * device array of WB_SOLVE_NOALLOC_BYTES that the library holds on each device it runs on until
 * the process ends; the call allocates nothing. Problems whose wb_solve_workspace_size exceeds
 * WB_SOLVE_NOALLOC_BYTES are refused with WB_OOM. Every call on one device shares that device's
 * array: calls that overlap on one device (two streams or two threads) race, so serialize them.
 * Arguments and outputs are wb_solve's, less the off-diagonal slots (Li and Laurent layers). */
#define WB_SOLVE_NOALLOC_BYTES (64u << 20)
int wb_solve_noalloc(wb_ctx* ctx, int P, int L, int hu, int hv, int mode, const double* pts,
                     const double* thick, const int* conv_idx, const double* epx, const double* epy,
                     const double* epz_i, double* eff, double* amp, double* pe, int* flags,
                     void* stream);

/* The layer modes of a batch, the first half of wb_solve: for each (point, layer) pair b = p L + l
 * the raw layer eigenmodes of rcwa/_engine/solve_2d.py:2303-2581 layer_modes_2d (W = [Ex; Ey]
 * eigenvectors of Omega2 with unit columns, V = Omega_R W / q, q by the SMM root rule) or of
 * rcwa/_engine/solve_1d.py:151-253 layer_modes_1d (TE: V = W q; TM: V = epx^-1 W q), before the
 * s/p rotation of solve_2d.py:2584-2616. W, V: P L x n x n complex; q: P L x n complex. flags
 * P x L (or NULL): WB_FLAG_NOCONVERGE, the pair screen's bits and, with rebuild = 1 on the 2D
 * path, WB_FLAG_REBUILT / WB_FLAG_DEFECTIVE (the grazing clusters re-derived as
 * wb_solve_rebuild does, n <= 2048). The modes depend on the point's wavelength, angles, frame
 * and half-space media and on the layer's conv matrices, not on its thickness. */
size_t wb_layer_modes_workspace_size(int P, int L, int hu, int hv, int mode, int rebuild);
int wb_layer_modes(wb_ctx* ctx, int P, int L, int hu, int hv, int mode, const double* pts,
                   const int* conv_idx, const double* epx, const double* epy, const double* epz_i,
                   const double* epxy, const double* epyx, double* W, double* V, double* q, int* flags, int rebuild, void* workspace,
                   size_t workspace_bytes, void* stream);

/* The second half of wb_solve: the layer S-matrices, the cascade and the efficiencies from modes
 * of wb_layer_modes (rcwa/_engine/scattering.py:1185-1324 _smm_sp from its per-layer modes).
 * Pair b = p L + l takes the modes at index modes_idx[b] of W, V, q (NULL: index b), so modes
 * computed once serve runs that change only thicknesses, and one wb_layer_modes call can serve
 * several calls with different layer counts. flags P x L (or NULL) report this call's singular
 * factorizations only; OR them with the modes' flags. smat is NULL or wb_solve_scattering's
 * output. The workspace is wb_solve_workspace_size's. Results equal wb_solve's to the bit. */
int wb_solve_from_modes(wb_ctx* ctx, int P, int L, int hu, int hv, int mode, const double* pts,
                        const double* thick, const double* W, const double* V, const double* q,
                        const int* modes_idx, double* eff, double* amp, double* pe, int* flags,
                        double* smat, void* workspace, size_t workspace_bytes, void* stream);

/* wb_solve_from_modes over points whose layer counts differ, as one cascade: each point runs the
 * reference's per-problem cascade over its own layers (rcwa/_engine/scattering.py:1185-1324
 * _smm_sp, called once per problem), here batched so that cascade step l takes layer l of every
 * point that has it. layers (host memory, P entries) are the points' layer counts, each in
 * 1..64 and non-increasing (order the points so). Pairs are numbered layer-major: with P_l the
 * number of points with more than l layers and base_l = P_0 + ... + P_{l-1}, layer l of point p
 * is pair base_l + p. thick (B doubles, B = the sum of the layer counts), modes_idx (B ints, or
 * NULL: index b) and flags (B ints, or NULL) follow that numbering; pts, eff, amp and pe are per
 * point as in wb_solve. No scattering output. Results equal wb_solve_from_modes' for each point
 * within rounding (the same operations per point in batches of other sizes). The workspace is
 * wb_solve_ragged_workspace_size's. */
size_t wb_solve_ragged_workspace_size(int P, const int* layers, int hu, int hv, int mode);
int wb_solve_from_modes_ragged(wb_ctx* ctx, int P, const int* layers, int hu, int hv, int mode,
                               const double* pts, const double* thick, const double* W,
                               const double* V, const double* q, const int* modes_idx, double* eff,
                               double* amp, double* pe, int* flags, void* workspace,
                               size_t workspace_bytes, void* stream);

/* wb_solve (2D path, WB_MODE_2D) with every grazing layer's mode cluster re-derived from the
 * coupled first-order problem, as the reference always does (rcwa/_engine/solve_2d.py:118-228
 * _rebuild_grazing_cluster, entered at :2463-2485). Meant for the points whose wb_solve flags
 * carry WB_FLAG_GRAZING (a rare condition: an order on a layer's light line); the extra work is
 * one more n x n eigensolve per (point, layer) pair of the call. smat is NULL or the output of
 * wb_solve_scattering. Layers up to n = 2048 and clusters up to 8 modes are rebuilt. The
 * workspace is wb_solve_rebuild_workspace_size's. */
size_t wb_solve_rebuild_workspace_size(int P, int L, int hu, int hv);
int wb_solve_rebuild(wb_ctx* ctx, int P, int L, int hu, int hv, const double* pts,
                     const double* thick, const int* conv_idx, const double* epx, const double* epy,
                     const double* epz_i, const double* epxy, const double* epyx, double* eff, double* amp, double* pe, int* flags,
                     double* smat, void* workspace, size_t workspace_bytes, void* stream);

/* Field state: wb_solve_from_modes plus each layer's mode amplitudes for the point's own
 * incidence, the state rcwa/_engine/scattering.py:236-333 (_field_state) leaves for field
 * reconstruction (rcwa/_engine/fields.py). amps: P L x 2 x n complex, pair b = p L + l holding
 * the down-going amplitudes anchored at the layer's top face, then the up-going ones anchored at
 * its bottom face, in the basis of the modes (W, V, q) it was given: the field in the layer at
 * height z above its bottom face is E_t = W (a_up + a_down), with a_up = exp(-k0 q z) up and
 * a_down = exp(k0 q (z - d)) down. The incidence is unit E at the cell centre (a_s s + a_p p on
 * the 2D path; the s or p wave of the incident order on the 1D ones). The other outputs are
 * wb_solve_from_modes'. The workspace is wb_field_state_workspace_size's. */
size_t wb_field_state_workspace_size(int P, int L, int hu, int hv, int mode);
int wb_field_state(wb_ctx* ctx, int P, int L, int hu, int hv, int mode, const double* pts,
                   const double* thick, const double* W, const double* V, const double* q,
                   const int* modes_idx, double* amps, double* eff, double* amp, double* pe,
                   int* flags, void* workspace, size_t workspace_bytes, void* stream);

/* The six field components (Ex, Ey, Ez, Hx, Hy, Hz, H in units of E) of P points x L layers at
 * nz heights per layer, public frame (+z up, x and y from the cell centre): the expansion of
 * rcwa/_engine/field_result.py:109-114 over the components of rcwa/_engine/fields.py:503-814, in
 * one batched launch sequence. out (P x L nz x nxy x 6 complex, device) = scale[c] x component c
 * at local heights z[l nz + j] (device, L x nz, measured up from layer l's bottom face, within
 * its thickness) and points xy (device, nxy x, y pairs). Pair b = p L + l: pts and thick as
 * wb_solve's; its modes at modes_idx[b] (NULL: b) of W, V (n x n) and q (n) from
 * wb_layer_modes, its amplitudes at b of amps (2n) from wb_field_state, its inverse-rule matrix
 * at conv_idx[b] (NULL: b) of epz_i (N x N; not read on the TE path); scale: host array of 6
 * complex factors (1/Z0 on H gives A/m for E in V/m). The workspace is
 * wb_field_sample_workspace_size's. */
size_t wb_field_sample_workspace_size(int P, int L, int hu, int hv, int mode, int nz, int nxy);
int wb_field_sample(wb_ctx* ctx, int P, int L, int hu, int hv, int mode, const double* pts,
                    const double* thick, const double* W, const double* V, const double* q,
                    const int* modes_idx, const double* epz_i, const int* conv_idx,
                    const double* amps, int nz, const double* z, int nxy, const double* xy,
                    const double* scale, double* out, void* workspace, size_t workspace_bytes,
                    void* stream);

/* wb_solve plus the stack's scattering matrix per point: smat is P x 4 x n x n complex,
 * n = 2N on the 2D path (s amplitudes over p port variables) and N on the 1D paths, blocks in
 * the order (S11, S12, S21, S22), incidence-relative for top incidence: S11 reflects back into
 * the top medium, S21 transmits top to bottom, S12 bottom to top, S22 reflects into the bottom
 * medium. Every port amplitude is referenced at the cell centre; the p port variable is the p
 * wave's magnetic field (S11 e = -r_p, S21 e = t_p n_bot / n_top on the specular column e).
 * The workspace is wb_solve_workspace_size's. */
int wb_solve_scattering(wb_ctx* ctx, int P, int L, int hu, int hv, int mode, const double* pts,
                        const double* thick, const int* conv_idx, const double* epx,
                        const double* epy, const double* epz_i, const double* epxy,
                        const double* epyx, double* eff, double* amp, double* pe,
                        int* flags, double* smat, void* workspace, size_t workspace_bytes,
                        void* stream);

/* C[b] = alpha op(A[b]) op(B[b]) + beta C[b]; op 0 = as stored, 1 = conjugate transpose (then A
 * is stored k x m, B n x k). */
int wb_zgemm_batched_ops(wb_ctx* ctx, int opa, int opb, int m, int n, int k, const double* alpha,
                         const double* A, int lda, long long strideA, const double* B, int ldb,
                         long long strideB, const double* beta, double* C, int ldc,
                         long long strideC, int batch, void* stream);

/* op 0: solve A X = B; op 1: solve A^H X = B; both from wb_zgetrf_batched's factors of A. */
int wb_zgetrs_batched_ops(wb_ctx* ctx, int op, int n, int nrhs, const double* LU, int lda,
                          long long strideA, const int* ipiv, double* B, int ldb, long long strideB,
                          int batch, void* stream);

/* Gradients (2D path, WB_MODE_2D).
 * wb_solve_keep is wb_solve that also fills `checkpoint` (wb_solve_checkpoint_size bytes) with
 * what the reverse pass needs: per (point, layer) the layer eigenvectors, the inverses of the
 * rotated W and V, and the layer S11; per point the transmitted columns of the back-sweep.
 * The workspace of that call then holds the rest (the per-layer S12, the cascade's LU factors
 * and reflections) and must not be reused until wb_solve_adjoint has run.
 * wb_solve_adjoint, given the cotangent eff_bar (P x 3 x 4 x N, the layout of eff) of the
 * efficiencies of that solve, writes
 *   d_epx, d_epy, d_epz   n_conv x N x N complex: the cotangents (d/dRe + i d/dIm of a real
 *                         objective) of the conv matrices, summed over the (point, layer)
 *                         pairs that use each slot
 *   d_epxy, d_epyx        likewise for the off-diagonal blocks when the solve had them (an
 *                         NVF stack: epxy and epyx both given to wb_solve_keep and here;
 *                         both NULL otherwise, and then d_epxy, d_epyx may be NULL too)
 *   d_thick               P x L doubles: d/d thickness of each (point, layer).
 * The eigenbasis is held fixed and the modal constants vary as a matrix in it (the generator
 * rule), which gives the exact derivative of the efficiencies. */
size_t wb_solve_checkpoint_size(int P, int L, int hu, int hv, int mode);
int wb_solve_keep(wb_ctx* ctx, int P, int L, int hu, int hv, int mode, const double* pts,
                  const double* thick, const int* conv_idx, const double* epx, const double* epy,
                  const double* epz_i, const double* epxy, const double* epyx, double* eff,
                  double* amp, double* pe, int* flags, void* workspace, size_t workspace_bytes,
                  void* checkpoint, size_t checkpoint_bytes, void* stream);
size_t wb_solve_adjoint_workspace_size(int P, int L, int hu, int hv, int mode);
int wb_solve_adjoint(wb_ctx* ctx, int P, int L, int hu, int hv, int mode, const double* pts,
                     const double* thick, const int* conv_idx, int n_conv, const double* epx,
                     const double* epy, const double* epz_i, const double* epxy,
                     const double* epyx, const void* solve_workspace, const void* checkpoint,
                     const double* eff_bar, double* d_epx, double* d_epy, double* d_epz,
                     double* d_epxy, double* d_epyx, double* d_thick, void* workspace,
                     size_t workspace_bytes, void* stream);

/* Cotangent of the raster eps (L, Ny, Nx) from those of wb_assemble_li's outputs; epx and epz_i
 * are that call's outputs for the same eps. hu, hv <= 15. */
size_t wb_assemble_li_adjoint_workspace_size(int L, int Ny, int Nx, int hu, int hv);
int wb_assemble_li_adjoint(wb_ctx* ctx, const double* eps, int L, int Ny, int Nx, int hu, int hv,
                           const double* epx, const double* epz_i, const double* d_epx,
                           const double* d_epy, const double* d_epz, double* d_eps,
                           void* workspace, size_t workspace_bytes, void* stream);

/* Cotangent of the raster eps (L, Ny, Nx) from those of wb_assemble_laurent's outputs; epz_i is
 * that call's output for the same eps. hu, hv <= 15. */
size_t wb_assemble_laurent_adjoint_workspace_size(int L, int Ny, int Nx, int hu, int hv);
int wb_assemble_laurent_adjoint(wb_ctx* ctx, const double* eps, int L, int Ny, int Nx, int hu, int hv,
                                const double* epz_i, const double* d_epx, const double* d_epy,
                                const double* d_epz, double* d_eps, void* workspace,
                                size_t workspace_bytes, void* stream);

/* Cotangent of the raster eps (L, Ny, Nx) from those of wb_assemble_nvf's outputs (floor 0):
 * d_off is the sum of the cotangents of the two off-diagonal blocks (epxy = epyx). The normal
 * field is held fixed -- it is the geometry of the interfaces, which an infinitesimal change of
 * the values does not move (the reference's gradient treats it so: tests/test_frozen_geometry.py
 * :644) -- and only passive layers (Re eps > 0, Im eps >= 0 everywhere, where the floor of
 * factorization.py:1968-2160 is inactive) are differentiated; the caller refuses the others.
 * The forward is rebuilt inside (its workspace is part of this one). hu, hv <= 15. */
size_t wb_assemble_nvf_adjoint_workspace_size(int L, int Ny, int Nx, int hu, int hv);
int wb_assemble_nvf_adjoint(wb_ctx* ctx, const double* eps, int L, int Ny, int Nx, int hu, int hv,
                            const double* lattice, const double* d_epx, const double* d_epy,
                            const double* d_epz, const double* d_off, double* d_eps,
                            void* workspace, size_t workspace_bytes, void* stream);

/* Per-kernel and per-stage timing (no reference counterpart: measurement support).
 * wb_timing_enable(1) brackets the launches of the kernels below by hipEvent records on their
 * stream; wb_timing_enable(2) instead times whole stages of a solve: WB_T_STAGE_EIG (the layer
 * eigensolves, 100 n^3 flops each), WB_T_STAGE_SMAT (the rest of wb_solve: Omega^2, V, the
 * layer S-matrices and the cascade, 84 n^3 per layer-point + 96 (L+1) n^3 per point) and
 * WB_T_ASSEMBLE (the assembly gathers, timed in both modes); 0 turns timing off.
 * wb_timing_read returns the summed milliseconds, launch count and algorithmic real FP64 flops
 * (GEMM 8mnk; eigensolver stages from EIG(n) = 100 n^3) since the last reset. */
enum {
  WB_T_ZGEMM = 0, WB_T_GETRF_PANEL, WB_T_LASWP_TRSM, WB_T_TRSM_DIAG, WB_T_GEHRD, WB_T_UNGHR,
  WB_T_HQR, WB_T_TREVC, WB_T_COLNORM, WB_T_ASSEMBLE, WB_T_STAGE_EIG, WB_T_STAGE_SMAT, WB_T_COUNT
};
int wb_timing_enable(int on);
int wb_timing_read(int id, double* total_ms, long long* count, double* flops, int reset);

/* GEMM tile-shape autotuning (no reference counterpart: the reference calls BLAS, whose tile
 * choice is internal). Every batched complex GEMM of the engine picks one of
 * wb_gemm_tuning_configs() tile configurations per shape (op(A), op(B), m, n, k, batch); all
 * configurations accumulate in the same order and give bitwise identical results, so tuning
 * changes speed only. Untuned shapes use configuration 0 (64 x 64).
 *   wb_gemm_tuning_record(1)   subsequent launches record their shapes (0 stops recording);
 *   wb_gemm_autotune_size()    scratch bytes the recorded, untuned shapes need;
 *   wb_gemm_autotune(ctx, ...) times every configuration on each recorded, untuned shape on the
 *                              caller's scratch (zeroed) and keeps the fastest; it waits for the
 *                              stream, so call it outside any solve or graph capture. Returns
 *                              the number of shapes tuned (>= 0) or a negative wb_status;
 *   wb_gemm_tuning_table(out, cap) the table as text lines "opa opb m n k batch config",
 *                              returns the bytes needed (out may be NULL);
 *   wb_gemm_tuning_load(text)  merges such lines, returns how many were taken;
 *   wb_gemm_tuning_set(...)    forces one shape's configuration (config < 0 removes it);
 *   wb_gemm_tuning_clear()     forgets the table and the recorded shapes.
 * The table is per process and shared by all contexts. */
int wb_gemm_tuning_configs(void);
const char* wb_gemm_tuning_config_name(int config);
int wb_gemm_tuning_record(int on);
size_t wb_gemm_autotune_size(void);
int wb_gemm_autotune(wb_ctx* ctx, int reps, void* scratch, size_t scratch_bytes, void* stream);
size_t wb_gemm_tuning_table(char* out, size_t cap);
int wb_gemm_tuning_load(const char* text);
int wb_gemm_tuning_set(int opa, int opb, int m, int n, int k, int batch, int config);
void wb_gemm_tuning_clear(void);

#ifdef __cplusplus
}
#endif

#endif /* WAVEBIT_AMD_H */
