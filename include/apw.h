/*
 * apw.h -- C ABI of the MI355X-native plane-wave Davidson path.
 *
 * Replaces, for one k-point / one GPU, the device path of ABACUS's
 *   hsolver::DiagoDavid<T, DEVICE_GPU>          (source/module_hsolver/diago_david.h:21-84)
 * driving the HamiltPW operator chain
 *   Operator<T, DEVICE_GPU>::hPsi / act          (source/module_hamilt_general/operator.h:48-65)
 *   Ekinetic / Veff / Nonlocal <OperatorPW>      (source/module_hamilt_pw/hamilt_pwdft/operator_pw/)
 * and the device functors they call.
 *
 * Conventions
 *  - Every function returns APW_OK (0) or a negative APW_E* status; the message of the
 *    last failure on the calling thread is apw_last_error().  Nothing calls exit().
 *  - Complex arrays are interleaved (re, im) pairs: "c128" = 2 doubles, "c64" = 2 floats.
 *  - Matrices are column-major, as in the reference (psi: ld x nbands, band b at psi + b*ld).
 *  - Pointers documented "device" are HIP device pointers owned by the caller; "host"
 *    pointers are read during the call only.
 *  - All work of a context is ordered on the context's stream.  Device outputs are
 *    ready after apw_ctx_synchronize() (host outputs are ready on return).
 */
#ifndef APW_H
#define APW_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    APW_OK = 0,
    APW_EINVAL = -1,      /* bad argument / shape                        */
    APW_EHIP = -2,        /* HIP runtime error                           */
    APW_ENOMEM = -3,      /* device allocation failed                    */
    APW_EBREAKDOWN = -4,  /* Gram-Schmidt norm < 1e-12 (reference aborts) */
    APW_EEIG = -5,        /* subspace eigensolver did not converge        */
    APW_ECALLBACK = -6    /* a user callback returned non-zero            */
};

/* APW_FP32_MIXED_X32 (Davidson only; no counterpart in the reference): the FP32-mixed form
 * with the pre-phase's subspace products (cal_grad's residual block, cal_elem's new hcc rows)
 * also on the FP32 matrix cores (c128 operands rounded to FP32, FP64 sums); the
 * orthonormalisation, the restart vectors and the FP64 refinement stay FP64. */
enum { APW_FP64 = 0, APW_FP32_MIXED = 1, APW_FP32_MIXED_X32 = 2 };

typedef struct apw_ctx apw_ctx;
typedef struct apw_ham apw_ham;
typedef struct apw_david apw_david;

const char* apw_last_error(void);

/* Context: one device, one stream.  stream == NULL: the library creates its own
 * (default priority) stream.  No counterpart in the reference, which uses the
 * default stream and global hipBLAS handles (rocm/math_kernel_op.hip.cu:40,52). */
int apw_ctx_create(apw_ctx** out, int device, void* hip_stream);
int apw_ctx_destroy(apw_ctx* ctx);
int apw_ctx_synchronize(apw_ctx* ctx);
void* apw_ctx_stream(apw_ctx* ctx);

/* No counterpart in the reference.  Subspace matrices of the solvers on this context (and of
 * apw_dnevx_op / apw_dngvx_op) with n <= nmax are reduced to tridiagonal form by a single
 * workgroup with an ordinary launch instead of the cooperative multi-workgroup reduction, so
 * many small solves on different streams (batched k-points or cells) run at the same time;
 * 0 (the default) never does.  The eigenpairs are those of the cooperative form. */
int apw_ctx_set_eig_single_wg(apw_ctx* ctx, int nmax);
/* Kernel profiler (no counterpart in the reference, whose ModuleBase::timer is host-side,
 * timer.cpp:71-130): when on, the library's kernels are bracketed by HIP events on the
 * context stream and accumulated per kernel name with their algorithmic HBM bytes and
 * flops per launch.  set_profiling clears the table.  profile_count resolves the events
 * (waits for them) and returns the number of entries. */
int apw_ctx_set_profiling(apw_ctx* ctx, int on);
int apw_ctx_profile_count(apw_ctx* ctx);
int apw_ctx_profile_entry(apw_ctx* ctx, int i, char* name, int name_cap, long long* launches, double* ms,
                          double* bytes, double* flops);

/* ------------------------------------------------------------------------------
 * Hamiltonian of one k-point (HamiltPW + PW_Basis_K + pseudopot_cell_vnl data).
 * Host arrays, uploaded once:
 *   ig2ixyz[npw]     iz + iy*nz + ix*ny*nz     (PW_Basis_K::ig2ixyz_k, pw_basis_k.cpp:335)
 *   gk2[npw]         |k+G|^2 in tpiba^2 units   (pw_basis_k.cpp:247)
 *   veff[nx*ny*nz]   local potential, z fastest (Veff::act, veff_pw.cpp:63)
 *   vkb              c128, nkb rows of npwx (row-major nkb x npwx = col-major npwx x nkb,
 *                    VNL_in_pw.cpp:226), projectors ordered type -> atom -> ih
 *   deeq[nat*nhm*nhm] real, spin 0, deeq[(iat*nhm + x)*nhm + k] (nonlocal_op.cpp:33-37)
 *   na[ntype], nh[ntype]
 * Replaces HamiltPW construction (hamilt_pw.cpp:22-114) + Nonlocal::init/getvnl.
 * ------------------------------------------------------------------------------ */
typedef struct apw_ham_desc {
    int nx, ny, nz;
    int npw, npwx;
    const int* ig2ixyz;
    const double* gk2;
    double tpiba2;
    const double* veff;
    int nkb;
    const double* vkb;
    int ntype;
    const int* na;
    const int* nh;
    int nhm;
    const double* deeq;
    int precision;      /* APW_FP64 or APW_FP32_MIXED: arithmetic of FFT/Veff/nonlocal */
} apw_ham_desc;

int apw_ham_create(apw_ctx* ctx, const apw_ham_desc* desc, apw_ham** out);
int apw_ham_destroy(apw_ham* ham);
/* Switch the H|psi> arithmetic (APW_FP64 / APW_FP32_MIXED); psi and hpsi stay c128. */
int apw_ham_set_precision(apw_ham* ham, int precision);

/* Replace the projectors with device-resident ones (e.g. from apw_getvnl): vkb device c128,
 * nkb rows of ldvkb (>= npw), the first npw entries of each row are used.  This is where
 * pseudopot_cell_vnl::getvnl(ctx, ik, vkb) (VNL_in_pw.cpp:412-524) hands ppcell->vkb to the
 * nonlocal operator (nonlocal_pw.cpp:58-64). */
int apw_ham_set_vkb(apw_ham* ham, const double* vkb, int ldvkb);

/* Operator::hPsi (operator.cpp:39-88): hpsi[:, b] = H psi[:, b] for b < nbands
 * (overwrite; the reference zeroes then accumulates every act()).  Device c128. */
int apw_hpsi(apw_ham* ham, const double* psi, int ldpsi, double* hpsi, int ldhpsi, int nbands);

/* Operator::act of each operator (ACCUMULATE into hpsi), nbasis = leading dimension:
 * Ekinetic  ekinetic_pw.cpp:34-58; Veff veff_pw.cpp:38-105; Nonlocal nonlocal_pw.cpp:211-304 */
int apw_act_ekinetic(apw_ham* ham, int nbands, int nbasis, const double* psi, double* hpsi);
int apw_act_veff(apw_ham* ham, int nbands, int nbasis, const double* psi, double* hpsi);
int apw_act_nonlocal(apw_ham* ham, int nbands, int nbasis, const double* psi, double* hpsi);

/* PW_Basis_K::recip_to_real / real_to_recip, GPU forms (pw_transform_k.cpp:336-476):
 * recip_to_real: box = FFT^+ (scatter(in)); out = box, or out += factor*box if add.
 * real_to_recip: out[ig] = box'[ig]/nxyz, or out[ig] += factor/nxyz * box'[ig] if add,
 *                box' = FFT^- (in).   in/out device c128 (box: nx*ny*nz). */
int apw_recip_to_real(apw_ham* ham, const double* in, double* out, int add, double factor);
int apw_real_to_recip(apw_ham* ham, const double* in, double* out, int add, double factor);

/* ------------------------------------------------------------------------------
 * Charge density from the bands (ElecStatePW::psiToRho / rhoBandK, nspin 1,
 * elecstate_pw.cpp:142-214): rho(r) = sum_b wg[b] / omega * |psi_b(r)|^2, psi_b(r) the
 * unnormalised backward FFT of band b scattered into the box (recip_to_real).
 * psi: device c128, band b at psi + b*ldpsi (ldpsi >= npw); wg: HOST double[nbands];
 * rho: device double[nx*ny*nz], overwritten (add = 0) or accumulated (add != 0).
 * Bands are summed in order (deterministic).  Synchronous. */
int apw_psi_to_rho(apw_ham* ham, const double* psi, int ldpsi, int nbands, const double* wg,
                   double omega, double* rho, int add);

/* ------------------------------------------------------------------------------
 * Projector generation for one k-point (pseudopot_cell_vnl::getvnl, GPU branch,
 * VNL_in_pw.cpp:412-524: Ylm_Real, get_sk, cal_vnl_op).
 * ------------------------------------------------------------------------------ */
/* cal_ylm_real_op<double, DEVICE_GPU> (module_base/kernels/math_op.h:12-39): real
 * spherical harmonics of ng points g (device, ng x 3) into ylm ((lmax+1)^2 x ng), the
 * Legendre values into p ((lmax+1)^2 x ng, may be NULL).  lmax <= 6. */
int apw_cal_ylm_real_op(apw_ctx* ctx, int ng, int lmax, double sqrt2, double pi, double pi_half,
                        double four_pi, double sqrt_inverse_four_pi, const double* g, double* p,
                        double* ylm);
/* cal_sk_op<double, DEVICE_GPU> (kernels/wf_op.h:10-58): structure factor sk (nat x npw
 * c128) from the eigts tables.  All arrays device. */
int apw_cal_sk_op(apw_ctx* ctx, int ik, int ntype, int nx, int ny, int nz, int rho_nx, int rho_ny,
                  int rho_nz, int npw, int npwx, int fftny, int eigts1_nc, int eigts2_nc,
                  int eigts3_nc, const int* atom_na, const int* igl2isz, const int* is2fftixy,
                  double two_pi, const double* kvec_c, const double* atom_tau, const double* eigts1,
                  const double* eigts2, const double* eigts3, double* sk);
/* cal_vnl_op<double, DEVICE_GPU> (kernels/vnl_op.h:10-58): vkb_in (nkb x npwx c128) from
 * gk (npw x 3), ylm, the radial table tab (ntype x tab_2 x tab_3) and sk; vkb1 (nhm x npw)
 * is scratch, as in the reference.  NEG_IMAG_UNIT passed as (re, im).  All arrays device.
 * Fails (APW_EINVAL) when |k+G| tpiba / dq reaches past the table. */
int apw_cal_vnl_op(apw_ctx* ctx, int ntype, int npw, int npwx, int nhm, int tab_2, int tab_3,
                   const int* atom_na, const int* atom_nb, const int* atom_nh, double dq,
                   double tpiba, double neg_imag_re, double neg_imag_im, const double* gk,
                   const double* ylm, const double* indv, const double* nhtol,
                   const double* nhtolm, const double* tab, double* vkb1, const double* sk,
                   double* vkb_in);
/* getvnl(ctx, ik, vkb_in) in one pass: the three functors above fused, the structure
 * factor as exp(-2 pi i (k+G).tau).  Sizes and atom_na/nb/nh, indv/nhtol/nhtolm (ntype x
 * nhm) and tau (nat x 3, cartesian, lat0 units) are HOST arrays; gk (k+G cartesian, 2 pi /
 * lat0 units, npw x 3), tab and vkb are device.  ntype <= 16, nhm <= 32, l <= 3. */
int apw_getvnl(apw_ctx* ctx, int ntype, int npw, int npwx, int nhm, int tab_2, int tab_3,
               const int* atom_na, const int* atom_nb, const int* atom_nh, double dq, double tpiba,
               const double* gk, const double* tau, const double* indv, const double* nhtol,
               const double* nhtolm, const double* tab, double* vkb);

/* ------------------------------------------------------------------------------
 * Plane-wave wavefunction files of one k-point, npol 1 (the reference's file layouts).
 * ------------------------------------------------------------------------------ */
/* ModuleIO::write_wfc_pw (module_io/write_wfc_pw.cpp:13-283): filename ending ".dat"
 * (binary records, out_wfc_pw 2) or ".txt" (text, out_wfc_pw 1), created or truncated.
 * ik_global: 0-based k index among nkstot; kvec_c[3] (2 pi / lat0), weight, ecutwfc, lat0,
 * tpiba and G (3 x 3 reciprocal vectors, row-major, 2 pi / lat0) go into the header;
 * ig2ixyz HOST int[npw] (iz + iy*nz + ix*ny*nz of each plane wave, basis order);
 * psi DEVICE c128, band b at psi + b*ldpsi. */
int apw_write_wfc_pw(apw_ctx* ctx, const char* filename, int ik_global, int nkstot,
                     const double* kvec_c, double weight, int nbands, double ecutwfc, double lat0,
                     double tpiba, const double* G, int nx, int ny, int nz, int npw,
                     const int* ig2ixyz, const double* psi, int ldpsi);
/* ModuleIO::read_wfc_pw (module_io/read_wfc_pw.cpp:12-341): ".dat" files only (the
 * reference reads nothing from ".txt").  The header must match ik + 1, nkstot, npw, nbands,
 * kvec_c, lat0, tpiba and G exactly (the reference's checks); the file's plane waves are
 * mapped onto this basis through their Miller indices.  wfc DEVICE c128 (ld ldwfc) out. */
int apw_read_wfc_pw(apw_ctx* ctx, const char* filename, int ik, int nkstot, int nbands,
                    const double* kvec_c, double lat0, double tpiba, const double* G, int nx,
                    int ny, int nz, int npw, const int* ig2ixyz, double* wfc, int ldwfc);

/* ------------------------------------------------------------------------------
 * Block Davidson (DiagoDavid, diago_david.cpp).
 * ------------------------------------------------------------------------------ */
/* DiagoDavid constructor (diago_david.cpp:34-121): precondition is HOST (len >= dim). */
int apw_david_create(apw_ctx* ctx, const double* precondition, int nband, int dim,
                     int david_ndim, apw_david** out);
int apw_david_destroy(apw_david* dav);

/* HPsiFunc (diago_david.h:56): HX = H * X[:, id_start..id_end], X = head of the block
 * (no offset, ld = dim), HX = exact destination.  Device c128.  Return 0 on success. */
typedef int (*apw_hpsi_cb)(void* user, const double* X, double* HX, int nvec, int dim,
                           int id_start, int id_end);
/* SPsiFunc (diago_david.h:75): SX = S X for nbands vectors of length npw (ld nrow). */
typedef int (*apw_spsi_cb)(void* user, const double* X, double* SX, int nrow, int npw, int nbands);

typedef struct apw_diag_stats {
    int iterations;        /* return value of DiagoDavid::diag: sum of dav_iter */
    int ntry;              /* diag_once calls                                    */
    int notconv;           /* bands unconverged at exit                          */
    int hpsi_calls;        /* H|psi> block applications                          */
    long long hpsi_bands;  /* bands pushed through H|psi>                        */
    int fp64_switch_iter;  /* FP32-mixed: iteration at which FP64 H|psi> took over (-1 = never) */
    double max_residual;   /* max ||H psi - e psi|| over bands at exit (FP64)    */
    double t_hpsi_s;       /* seconds inside H|psi> (device-event timed)         */
    double t_total_s;      /* seconds inside diag()                              */
} apw_diag_stats;

/* DiagoDavid::diag (diago_david.cpp:1173-1200).
 *   ham != NULL: H|psi> applied by the library (S = I); hpsi_cb/spsi_cb ignored.
 *   ham == NULL: hpsi_cb required; spsi_cb NULL means S = I.
 *   psi: device c128, ldpsi x nband, start vectors in, eigenvectors out.
 *   eig: HOST double[nband] out.  precision: APW_FP64, APW_FP32_MIXED or
 *   APW_FP32_MIXED_X32 (the mixed forms with the library's H only).
 * Returns the total Davidson iteration count (>= 0) or a negative status. */
int apw_david_diag(apw_david* dav, apw_ham* ham, apw_hpsi_cb hpsi_cb, apw_spsi_cb spsi_cb,
                   void* user, int ldpsi, double* psi, double* eig, double david_diag_thr,
                   int david_maxiter, int ntry_max, int notconv_max, int precision,
                   apw_diag_stats* stats);

/* DiagoDavid<std::complex<float>, DEVICE_GPU>::diag (diago_david.cpp:1246; ABACUS's
 * precision=single, esolver.cpp:137-141): psi device c64 (ldpsi x nband) in/out, H applied
 * by a c64 callback (HPsiFunc semantics, X = head of a c64 copy of the basis, HX = exact c64
 * destination, ld = dim), S = I.  The Davidson arithmetic is the FP64 form's on c128 copies
 * of psi and the blocks; the callback's FP32 products bound the attainable accuracy.
 * eig: HOST double[nband].  Returns the iteration count or a negative status. */
typedef int (*apw_hpsi_cb_c64)(void* user, const float* X, float* HX, int nvec, int dim, int id_start,
                               int id_end);
int apw_david_diag_c64(apw_david* dav, apw_hpsi_cb_c64 hpsi_cb, void* user, int ldpsi, float* psi,
                       double* eig, double david_diag_thr, int david_maxiter, int ntry_max,
                       int notconv_max, apw_diag_stats* stats);

/* Batched form: nunits independent DiagoDavid::diag calls (k-points or cells, uneven sizes),
 * each unit with its own solver, Hamiltonian (library H|psi>), psi (device, ld ldpsi[u]),
 * eig (host) and stats (optional array).  Units created on different contexts run
 * concurrently (one host thread and one stream each); units sharing a context run in
 * turn.  status[u]: the unit's iteration count or negative status.  Returns APW_OK, or the
 * first failing unit's status with its message in apw_last_error().  No single reference
 * function corresponds: HSolverPW::solve loops diag over k-points (hsolver_pw.cpp:472-545);
 * this is the proposed apw_david_diag_batched of the C-ABI boundary. */
int apw_david_diag_batched(int nunits, apw_david* const* dav, apw_ham* const* ham,
                           const int* ldpsi, double* const* psi, double* const* eig,
                           double david_diag_thr, int david_maxiter, int ntry_max,
                           int notconv_max, int precision, apw_diag_stats* stats, int* status);

/* dngvx_op<complex<double>, DEVICE_GPU> (dngvd_op.h:67-81): lowest m eigenpairs of the
 * Hermitian-definite pencil A x = lambda B x (nstart x nstart, ld ldh, both triangles;
 * the CPU op calls zhegvx).  eigenvalue device double[m]; vcc device c128 (ld ldh),
 * B-orthonormal.  dngvd_op (dngvd_op.h:24-45, zhegvd): all nstart pairs. */
int apw_dngvx_op(apw_ctx* ctx, int nstart, int ldh, const double* hcc, const double* scc,
                 int m, double* eigenvalue, double* vcc);
int apw_dngvd_op(apw_ctx* ctx, int nstart, int ldh, const double* hcc, const double* scc,
                 double* eigenvalue, double* vcc);

/* ------------------------------------------------------------------------------
 * Davidson with pencil Rayleigh-Ritz (Diago_DavSubspace, diago_dav_subspace.cpp; the
 * "dav_subspace" method of HSolverPW, hsolver_pw.cpp:434-471).
 *   create: precondition HOST double[dim] (HSolverPW passes |k+G|^2 tpiba2 for this
 *   method, hsolver_pw.cpp:570-577), nband, dim, david_ndim (> 1, david_ndim*nband < dim),
 *   diag_thr (eigenvalue-change threshold), diag_nmax (iterations per diag_once),
 *   need_subspace (Rayleigh-Ritz of the start vectors first).
 *   diag: as apw_david_diag (FP64 only); is_occupied HOST int[nband] or NULL (all
 *   occupied; unoccupied bands use max(5 thr, 1e-5)); scf selects test_exit_cond's rule.
 * Returns the total iteration count or a negative status.
 * ------------------------------------------------------------------------------ */
typedef struct apw_dav_subspace apw_dav_subspace;
int apw_dav_subspace_create(apw_ctx* ctx, const double* precondition, int nband, int dim,
                            int david_ndim, double diag_thr, int diag_nmax, int need_subspace,
                            apw_dav_subspace** out);
int apw_dav_subspace_destroy(apw_dav_subspace* dav);
int apw_dav_subspace_diag(apw_dav_subspace* dav, apw_ham* ham, apw_hpsi_cb hpsi_cb, void* user,
                          int ldpsi, double* psi, double* eig, const int* is_occupied, int scf,
                          apw_diag_stats* stats);

/* ------------------------------------------------------------------------------
 * Device functors of the path (parity entry points; same argument meaning as the
 * reference functor named beside each).  Device pointers unless noted.
 * ------------------------------------------------------------------------------ */
/* ekinetic_pw_op<double, DEVICE_GPU> (kernels/ekinetic_op.h; cuda/ekinetic_op.cu:13-48) */
int apw_ekinetic_pw_op(apw_ctx* ctx, int nband, int npw, int max_npw, double tpiba2,
                       const double* gk2, double* hpsi, const double* psi);
/* veff_pw_op<double, DEVICE_GPU> (cuda/veff_op.cu:14-23): out[ir] *= in[ir] */
int apw_veff_pw_op(apw_ctx* ctx, int size, double* out, const double* in);
/* elecstate_pw_op<double, DEVICE_GPU> (elecstate_op.h:11-31, cuda/elecstate_op.cu):
 * rho[spin][ir] += w1 * |wfcr[ir]|^2.  rho: HOST array of device pointers (one per
 * spin component); wfcr device c128[nrxx]. */
int apw_elecstate_pw_op(apw_ctx* ctx, int spin, int nrxx, double w1, double* const* rho,
                        const double* wfcr);
/* The non-collinear overload (elecstate_op.cpp:27-81): rho[0] += w1(|a|^2 + |b|^2);
 * domag: rho[1] += 2w1 Re(conj(a)b), rho[2] += 2w1 Im(conj(a)b), rho[3] += w1(|a|^2 - |b|^2);
 * domag_z only: rho[1] = rho[2] = 0, rho[3] += w1(|a|^2 - |b|^2); neither: rho[1..3] = 0. */
int apw_elecstate_pw_op_nc(apw_ctx* ctx, int domag, int domag_z, int nrxx, double w1,
                           double* const* rho, const double* wfcr, const double* wfcr_another_spin);
/* nonlocal_pw_op<double, DEVICE_GPU> (cuda/nonlocal_op.cu:14-38,95); sum and iat are
 * HOST in/out counters advanced as the reference does. */
int apw_nonlocal_pw_op(apw_ctx* ctx, int l1, int l2, int l3, int* sum, int* iat, int spin,
                       int nkb, int deeq_x, int deeq_y, int deeq_z, const double* deeq,
                       double* ps, const double* becp);
/* set_3d_fft_box_op / set_recip_to_real_output_op / set_real_to_recip_output_op
 * (module_pw/kernels/pw_op.h:9-48, cuda/pw_op.cu:12-122) */
int apw_set_3d_fft_box_op(apw_ctx* ctx, int npwk, const int* box_index, const double* in,
                          double* out);
int apw_set_recip_to_real_output_op(apw_ctx* ctx, int nrxx, int add, double factor,
                                    const double* in, double* out);
int apw_set_real_to_recip_output_op(apw_ctx* ctx, int npw_k, int nxyz, int add, double factor,
                                    const int* box_index, const double* in, double* out);
/* FFT::fft3D_backward / fft3D_forward (fft.cpp:773-821): in-place, unnormalised,
 * backward = exponent +i.  Device c128 box of nx*ny*nz, z fastest.  dir: +1 / -1. */
int apw_fft3d(apw_ctx* ctx, int nx, int ny, int nz, double* box, int dir);
/* gemm_op<complex<double>, DEVICE_GPU> (math_kernel_op.h:86-120; column-major zgemm
 * semantics; trans in 'N','T','C').  alpha/beta HOST c128 (2 doubles each). */
int apw_zgemm(apw_ctx* ctx, char transa, char transb, int m, int n, int k,
              const double* alpha, const double* a, int lda, const double* b, int ldb,
              const double* beta, double* c, int ldc);
/* gemm_op<complex<float>, DEVICE_GPU> (math_kernel_op.h:86-120): as apw_zgemm on c64 data,
 * FP32 MFMA with FP32 accumulation.  alpha/beta HOST c64 (2 floats each). */
int apw_cgemm(apw_ctx* ctx, char transa, char transb, int m, int n, int k,
              const float* alpha, const float* a, int lda, const float* b, int ldb,
              const float* beta, float* c, int ldc);
/* No counterpart in the reference: apw_zgemm's c128 product computed on the FP32 matrix
 * cores (operands rounded to FP32, FP32 partial sums over 32-64 k folded into FP64 ones,
 * C in c128), for callers that accept an error of ~1e-7 relative to sum |a||b| for twice
 * ~1.5x the FP64 rate on the Davidson shapes (the c128 operands keep the FP64 loads).
 * The Davidson solvers do not use it (measured: see the FP32-mixed notes). */
int apw_zgemm_x32(apw_ctx* ctx, char transa, char transb, int m, int n, int k,
                  const double* alpha, const double* a, int lda, const double* b, int ldb,
                  const double* beta, double* c, int ldc);
/* No counterpart in the reference: apw_zgemm for callers that cannot allocate device memory
 * (code under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * Same product, plan and result as apw_zgemm, but its split-K partials go to a fixed 48 MiB
 * device array of the library (one per GPU, held until the process ends) instead of a
 * workspace it allocates; it calls no hipMalloc/hipFree.  A product whose split-K partials
 * would not fit is rejected with APW_EINVAL (the default plan fits every shape on GPUs of
 * up to 256 CUs).  Asynchronous on the context's stream (no synchronisation).  Calls on one
 * GPU share the array: calls that overlap in time on one GPU (two contexts, two streams)
 * race and must be serialised by the caller. */
int apw_zgemm_noalloc(apw_ctx* ctx, char transa, char transb, int m, int n, int k,
                      const double* alpha, const double* a, int lda, const double* b, int ldb,
                      const double* beta, double* c, int ldc);
/* Writes the (row, column) of the MFMA result each lane and accumulator register holds, for
 * v_mfma_f64_16x16x4f64 and v_mfma_f32_16x16x4f32 (4 x 64 x 4 doubles at out_dev).  No
 * counterpart in the reference: a check of the operand layout the GEMMs rely on. */
int apw_probe_mfma_layout(apw_ctx* ctx, double* out_dev);
/* ---- autotuner.  No counterpart in the reference (its GPU path calls vendor libraries with
 * fixed arguments).  A process-wide table of measured choices per shape: GEMM plans (tile
 * shape, workgroup-to-XCD mapping, split-K rule), the subspace eigensolver's workgroup count,
 * and each FFT grid's pass schedule (register four-step or LDS Stockham kernels per axis) and
 * bands per launch.  mode 0: fixed rules (the default); 1: the table, fixed rules on a miss;
 * 2: as 1, and a miss is measured where it occurs (every candidate timed on the calling
 * stream, the fastest recorded).  The table is saved to / loaded from a JSON file. */
int apw_tune_set_mode(int mode);
int apw_tune_load(const char* path);   /* returns the number of entries read */
int apw_tune_save(const char* path);   /* returns the number of entries written */
int apw_tune_clear(void);
int apw_tune_stats(long long* entries, long long* hits, long long* misses, long long* tuned);
int apw_tune_entry(int i, char* key, int key_cap, int* a, int* b, double* us);
/* dnevx_op<complex<double>, DEVICE_GPU> (dngvd_op.h:85-107): lowest nbands eigenpairs
 * of the Hermitian nstart x nstart matrix hcc (lower triangle, ld ldh).  eigenvalue:
 * device double[nbands]; vcc: device c128, ld ldh, eigenvector j in column j. */
int apw_dnevx_op(apw_ctx* ctx, int nstart, int ldh, const double* hcc, int nbands,
                 double* eigenvalue, double* vcc);

#ifdef __cplusplus
}
#endif
#endif /* APW_H */
