// One relaxed-Picard iteration of the plain (reference-structured) path enqueued by a
// single call, so the host issues one C call per iteration instead of one per kernel.
// The kernels are the library's own entry points, called in the order of the loop body
// of SpectralSolver.step (sgh/solvers/spectral.py:384-497).
#include <hip/hip_runtime.h>

#include "sgh_common.h"

namespace {

inline const void* at(const void* base, long long elems, int esize) {
  return static_cast<const char*>(base) + elems * esize;
}

// three axis contractions x, y, z (cell-centred volume transform): in -> w1 -> w2 -> out
int volume_transform(const SghPlainIter* d, const void* Mx, const void* My, const void* Mz,
                     const void* in, double scale, void* out, void* st) {
  const int nx = d->nx, ny = d->ny, nz = d->nz;
  int rc = sgh_axis_contract(d->dtype, d->xform, (long long)nz * ny, nx, 1, Mx, in, 1.0,
                             d->vol_w1, st);
  if (rc != SGH_OK) return rc;
  rc = sgh_axis_contract(d->dtype, d->xform, nz, ny, nx, My, d->vol_w1, 1.0, d->vol_w2, st);
  if (rc != SGH_OK) return rc;
  return sgh_axis_contract(d->dtype, d->xform, 1, nz, (long long)ny * nx, Mz, d->vol_w2, scale,
                           out, st);
}

}  // namespace

extern "C" {

size_t sgh_plain_iter_size(void) { return sizeof(SghPlainIter); }

int sgh_plain_iter_fork_init(SghPlainIter* d) {
  SGH_REQUIRE(d && !d->side_stream && !d->ev_fork && !d->ev_join, SGH_ERR_ARG,
              "sgh_plain_iter_fork_init: null descriptor or already initialised");
  hipStream_t s = nullptr;
  hipEvent_t a = nullptr, b = nullptr;
  hipError_t e = hipStreamCreateWithFlags(&s, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&a, hipEventDisableTiming);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&b, hipEventDisableTiming);
  if (e != hipSuccess) {
    if (b) hipEventDestroy(b);
    if (a) hipEventDestroy(a);
    if (s) hipStreamDestroy(s);
    return sgh::fail_hip(e, "sgh_plain_iter_fork_init");
  }
  d->side_stream = s;
  d->ev_fork = a;
  d->ev_join = b;
  return SGH_OK;
}

int sgh_plain_iter_fork_release(SghPlainIter* d) {
  SGH_REQUIRE(d, SGH_ERR_ARG, "sgh_plain_iter_fork_release: null descriptor");
  hipError_t e = hipSuccess;
  if (d->side_stream) e = hipStreamSynchronize(reinterpret_cast<hipStream_t>(d->side_stream));
  if (d->ev_join) hipEventDestroy(reinterpret_cast<hipEvent_t>(d->ev_join));
  if (d->ev_fork) hipEventDestroy(reinterpret_cast<hipEvent_t>(d->ev_fork));
  if (d->side_stream) hipStreamDestroy(reinterpret_cast<hipStream_t>(d->side_stream));
  d->side_stream = d->ev_fork = d->ev_join = nullptr;
  if (e != hipSuccess) return sgh::fail_hip(e, "sgh_plain_iter_fork_release");
  return SGH_OK;
}

int sgh_plain_iterate(const SghPlainIter* d, int parity, void* stream) {
  SGH_REQUIRE(d && (parity == 0 || parity == 1), SGH_ERR_ARG, "sgh_plain_iterate: bad args");
  SGH_REQUIRE(d->latent_mode >= 0 && d->latent_mode <= 2, SGH_ERR_ARG,
              "sgh_plain_iterate: latent_mode %d", d->latent_mode);
  SGH_REQUIRE(d->norms_host && d->norms && d->partials && d->counter, SGH_ERR_ARG,
              "sgh_plain_iterate: missing norm buffers");
  const int es = d->dtype == SGH_F64 ? 8 : 4;
  const int nx = d->nx, ny = d->ny, nz = d->nz;
  const long long S = (long long)nx * ny, N = S * nz;
  const void* a_old = d->a_buf[parity];
  void* a_new = d->a_buf[parity ^ 1];
  int rc;
#define SGH_TRY(expr)            \
  do {                           \
    rc = (expr);                 \
    if (rc != SGH_OK) return rc; \
  } while (0)
  // The surface chain and the latent source depend only on a_old (and the z-partials the
  // previous fused pass left), so with a side stream they run concurrently: the side
  // stream starts after everything already on `stream` (the previous iteration, whose
  // consumers of S_top and T_s are then done) and `stream` waits for it before the ETD1
  // update.
  const bool fork = d->side_stream && d->latent_mode != 0;
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  hipStream_t side = reinterpret_cast<hipStream_t>(d->side_stream);
  void* sst = fork ? d->side_stream : stream;
  if (fork) {
    hipError_t e = hipEventRecord(reinterpret_cast<hipEvent_t>(d->ev_fork), st);
    if (e == hipSuccess) e = hipStreamWaitEvent(side, reinterpret_cast<hipEvent_t>(d->ev_fork), 0);
    if (e != hipSuccess) return sgh::fail_hip(e, "sgh_plain_iterate fork");
  }
  // top-surface temperature of a_old: the z-partials left by the previous fused pass
  SGH_TRY(sgh_sum_parts(d->dtype, S, d->nchunks, d->zc, sst));
  SGH_TRY(sgh_sep2d_forcing(d->dtype, ny, nx, d->sByT, d->sBx, d->zc, 1.0, d->q_las, d->Pa,
                            d->T_boil, d->dH, d->R_v, d->T_liq, d->h_top, d->T0,
                            d->rescale ? d->prog : nullptr, d->kbar, d->T_s,
                            d->q_evap[parity], d->q_tmp, sst));
  SGH_TRY(sgh_sep2d(d->dtype, 1, ny, nx, d->sBy, d->sBxT, d->q_tmp, 1, d->dxdy, d->S_top,
                    sst));
  if (fork) {
    hipError_t e = hipEventRecord(reinterpret_cast<hipEvent_t>(d->ev_join), side);
    if (e != hipSuccess) return sgh::fail_hip(e, "sgh_plain_iterate join");
  }
  const void* Q_modes = nullptr;
  const void* lprog = d->rescale ? d->prog : nullptr;
  if (d->latent_mode == 2) {
    const int nxb = d->nxb, nyb = d->nyb, nzb = d->nzb;
    SGH_TRY(sgh_box_latent_source(d->dtype, d->xform, nx, ny, nz, a_old, nxb, nyb, nzb,
                                  at(d->fBxT, (long long)d->x_start * nx, es), nx,
                                  at(d->fByT, (long long)d->y_start * ny, es), ny, d->fBzT,
                                  d->T_prev_box, d->T_S, d->T_L, d->rho, d->L_f, d->dt, lprog,
                                  d->box_work, d->Q_latent[parity], stream));
    const long long work = (long long)nzb * (ny > nyb ? ny : nyb) * nx;
    void* w1 = d->box_work;
    void* w2 = static_cast<char*>(d->box_work) + work * es;
    SGH_TRY(sgh_axis_contract_strided(d->dtype, d->xform, (long long)nzb * nyb, nxb, nx, 1,
                                      at(d->fBx, d->x_start, es), d->ntx, d->Q_latent[parity],
                                      1.0, w1, stream));
    SGH_TRY(sgh_axis_contract_strided(d->dtype, d->xform, nzb, nyb, ny, nx,
                                      at(d->fBy, d->y_start, es), d->nty, w1, 1.0, w2, stream));
    SGH_TRY(sgh_axis_contract_strided(d->dtype, d->xform, 1, nzb, nz, S, d->fBz, nzb, w2,
                                      d->dV_fine, d->Q_modes[parity], stream));
    Q_modes = d->Q_modes[parity];
  } else if (d->latent_mode == 1) {
    SGH_TRY(volume_transform(d, d->vBxT, d->vByT, d->vBzT, a_old, 1.0, d->vol_T, stream));
    SGH_TRY(sgh_latent_source(d->dtype, N, d->vol_T, d->T_prev_full, d->T_S, d->T_L, d->rho,
                              d->L_f, d->dt, lprog, d->Q_latent[parity], stream));
    SGH_TRY(volume_transform(d, d->vBx, d->vBy, d->vBz, d->Q_latent[parity], d->dV,
                             d->Q_modes[parity], stream));
    Q_modes = d->Q_modes[parity];
  }
  if (fork) {
    hipError_t e = hipStreamWaitEvent(st, reinterpret_cast<hipEvent_t>(d->ev_join), 0);
    if (e != hipSuccess) return sgh::fail_hip(e, "sgh_plain_iterate join");
  }
  const void* wb = d->have_bottom ? d->w_bot : nullptr;
  const void* Sb = d->have_bottom ? d->S_bot_raw : nullptr;
  if (d->rescale) {
    SGH_TRY(sgh_etd1_update(d->dtype, nx, ny, nz, d->aK, d->KK, d->w_top, d->S_top, wb, Sb,
                            Q_modes, d->a_raw, stream));
    SGH_TRY(volume_transform(d, d->vBxT, d->vByT, d->vBzT, d->a_raw, 1.0, d->vol_T, stream));
    SGH_TRY(sgh_property_correction_source(d->dtype, nx, ny, nz, d->vol_T, d->T_prev_full,
                                           d->prog, d->kbar, d->abar, d->dt, d->dx, d->dy,
                                           d->dz, d->corr_work, d->corr_f, stream));
    SGH_TRY(volume_transform(d, d->vBx, d->vBy, d->vBz, d->corr_f, d->dV, d->C_modes, stream));
    SGH_TRY(sgh_picard_iter_fused_norms(d->dtype, nx, ny, nz, d->a_raw, d->KK, a_old, d->w_top,
                                        d->S_zero, nullptr, nullptr, d->C_modes, d->omega,
                                        d->w_top, d->nchunks, d->zc, a_new, d->partials,
                                        d->counter, d->norms, d->norms_host + 4 * parity,
                                        d->seq, stream));
  } else {
    SGH_TRY(sgh_picard_iter_fused_norms(d->dtype, nx, ny, nz, d->aK, d->KK, a_old, d->w_top,
                                        d->S_top, wb, Sb, Q_modes, d->omega, d->w_top,
                                        d->nchunks, d->zc, a_new, d->partials, d->counter,
                                        d->norms, d->norms_host + 4 * parity, d->seq, stream));
  }
#undef SGH_TRY
  return SGH_OK;
}

}  // extern "C"
