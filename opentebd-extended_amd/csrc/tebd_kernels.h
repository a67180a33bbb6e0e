// Host-side launchers of the gfx950 kernels (tebd_kernels.hip, tebd_theta_mixed.hip, tebd_jacobi.hip, tebd_qrp.hip,
// tebd_sketch.hip, tebd_canonical.hip, tebd_observables.hip, tebd_mlp.hip).
#pragma once
#include <hip/hip_runtime.h>
#include "tebd_device.h"

hipError_t launch_theta_gate(const BondTask* d_tasks, int ntasks, int d, int max_chil, int max_chir, hipStream_t st);
hipError_t launch_theta_gate_mfma(const BondTask* d_tasks, int ntasks, int d, int max_chil, int max_chir,
                                  hipStream_t st);
hipError_t launch_theta_gate_mixed(const BondTask* d_tasks, int ntasks, int d, int max_chil, int max_chir,
                                   hipStream_t st);
// Split-operand theta + gate (int8 matrix cores, exact integer accumulation, FP64
// combination): slices A and B into ws, then theta into W.
hipError_t launch_theta_gate_split(const BondTask* d_tasks, int ntasks, int d, int max_chil, int max_chir,
                                   const SplitWs& ws, hipStream_t st);
hipError_t launch_prepare_svd(const BondTask* d_tasks, int ntasks, hipStream_t st);
hipError_t launch_frobenius(const BondTask* d_tasks, int ntasks, int max_cpad, hipStream_t st, int keep_sum);
hipError_t launch_jacobi_begin(const BondTask* d_tasks, int ntasks, hipStream_t st);
hipError_t launch_jacobi_sweep_begin(const BondTask* d_tasks, int ntasks, int sweep, int* wl_count, hipStream_t st);
hipError_t launch_jacobi_round(const BondTask* d_tasks, int ntasks, int max_pairs, int sweep, int round,
                               double tol_rel, double tol_abs, int max_rows_w, int max_rows_v, int* wl_count,
                               int2* wl_items, int wl_cap, int inner_sweeps, hipStream_t st);
hipError_t launch_jacobi_rotate(const BondTask* d_tasks, int round, int max_rows_w, int max_rows_v, int* wl_count,
                                int2* wl_items, int wl_cap, hipStream_t st);
hipError_t launch_qrp_trailing(const BondTask* d_tasks, int ntasks, int max_rb, int max_cb, int j0, hipStream_t st);
hipError_t launch_qrp_begin(const BondTask* d_tasks, int ntasks, int max_cb, double tol_rel, hipStream_t st);
hipError_t launch_qrp_panel(const BondTask* d_tasks, int ntasks, int max_rb, int max_cb, int j0, hipStream_t st);
hipError_t launch_qrp_build_x(const BondTask* d_tasks, int ntasks, hipStream_t st);
hipError_t launch_qrp_vectors(const BondTask* d_tasks, int ntasks, int max_k, int max_r, hipStream_t st);
hipError_t launch_qrp_form_q(const BondTask* d_tasks, int ntasks, int max_r, hipStream_t st);
hipError_t launch_zdotc(const BondTask* d_a, const BondTask* d_b, int ntasks, double* d_out, hipStream_t st);
hipError_t launch_finalize(const BondTask* d_tasks, int ntasks, int max_cpad, double cutoff, hipStream_t st);
hipError_t launch_mlp_features(const BondTask* d_tasks, int ntasks, float* d_X, hipStream_t st);
hipError_t launch_mlp_predict(const float* d_params, const float* d_X, int n, float* d_out, hipStream_t st);
hipError_t launch_write_gammas(const BondTask* d_tasks, int ntasks, hipStream_t st);
// tile variants of the batched complex GEMM (chosen per shape by the autotuner)
enum ZgemmVariant {
  ZGEMM_VALU_64x64 = 0,
  ZGEMM_MFMA_64x64 = 1,
  ZGEMM_MFMA_128x64 = 2,
  ZGEMM_MFMA_64x96 = 3,
  ZGEMM_MFMA_64x64_K32 = 4,  // 64 x 64 tiles, k chunks of 32
  ZGEMM_MFMA_128x96_W8 = 5,  // 128 x 96 tiles on 8 waves
  ZGEMM_VARIANTS = 6
};
hipError_t launch_zgemm(const GemmTask* d_tasks, int ntasks, int max_m, int max_n, bool conj_a, hipStream_t st,
                        int variant);
hipError_t launch_build_sketch(const BondTask* d_tasks, const int* d_kprime, int ntasks, unsigned long long seed,
                               hipStream_t st);
hipError_t launch_probe_coeffs(const ProbeTask* d_tasks, int ntasks, int max_kn, unsigned long long seed,
                               hipStream_t st);
hipError_t launch_probe_norm(const ProbeTask* d_tasks, int ntasks, hipStream_t st);
hipError_t launch_col_norms(const BondTask* d_tasks, int ntasks, int max_cpad, hipStream_t st);
hipError_t launch_truncate(const BondTask* d_tasks, int ntasks, double cutoff, hipStream_t st);
hipError_t launch_scaled_copy(const zc* in, int ldi, zc* out, int ldo, int rows, int cols, const double* rowscale,
                              int row_mod, const double* colscale, int col_div, int conj_t, hipStream_t st);
hipError_t launch_perm_cols(const zc* src, int lds, zc* dst, int ldd, int rows, int ncols, const int* perm,
                            const double* s, int mode, double thr, hipStream_t st);
hipError_t launch_cf_xy(const zc* U1, const double* s1, const zc* U2, const double* s2, const double* lam1, zc* X,
                        zc* Y, int n, hipStream_t st);
hipError_t launch_cf_lambda(const double* s, int n, double* lam, int chi_max, hipStream_t st);
hipError_t launch_scale_vec(const double* lambda, const double* phase, int chi, int d, double* s, hipStream_t st);
hipError_t launch_bra_kernal(const zc* G, const double* lam1, const double* lam2, int chi_l, int d, int chi_r, zc* out,
                             hipStream_t st);
hipError_t launch_rho_reduce(const zc* G, const double* s, const zc* T, int chi_a, int chi_b, int d, zc* rho,
                             hipStream_t st);
hipError_t launch_rho1_trace(const zc* Theta, int chi, int d, zc* rho, hipStream_t st);
hipError_t launch_shrink_slab(const zc* src, int cl, int d, zc* dst, int ncl, int ncr, hipStream_t st);
hipError_t launch_zero_lambda_tail(double* lam, int nbonds, int chi, int chi_max, hipStream_t st);
hipError_t launch_nc_gather(const NcBlock* d_blocks, int nblocks, double* d_partial, hipStream_t st);
int nc_gather_split();
hipError_t launch_nc_vectors(const NcBlock* d_blocks, const BondTask* d_tasks, int nblocks, int max_k, hipStream_t st);
hipError_t launch_nc_gammas(const NcKept* d_kept, int nkept, double cut, hipStream_t st);
