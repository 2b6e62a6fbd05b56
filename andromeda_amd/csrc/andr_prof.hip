// Per-kernel event timing (see andr_prof.h), its C-ABI readers and those of the autotuner table.
#include "andr_prof.h"

#include <cstring>
#include <mutex>

static std::map<const andr_ctx *, ProfState> g_prof;
static std::mutex g_prof_mu;

ProfState &andr_prof(andr_ctx *ctx) {
  std::lock_guard<std::mutex> lk(g_prof_mu);
  return g_prof[ctx];
}

static hipEvent_t take_event(ProfState &P) {
  if (!P.pool.empty()) {
    hipEvent_t e = P.pool.back();
    P.pool.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  (void)hipEventCreate(&e);
  return e;
}

int andr_prof_begin(andr_ctx *ctx, const char *name, double bytes, double flops) {
  ProfState &P = andr_prof(ctx);
  if (!P.on) return -1;
  ProfRec r;
  r.name = name;
  r.a = take_event(P);
  r.b = take_event(P);
  r.bytes = bytes;
  r.flops = flops;
  (void)hipEventRecord(r.a, ctx->stream);
  P.pending.push_back(r);
  return (int)P.pending.size() - 1;
}

void andr_prof_end(andr_ctx *ctx, int token) {
  if (token < 0) return;
  ProfState &P = andr_prof(ctx);
  (void)hipEventRecord(P.pending[token].b, ctx->stream);
  if (P.pending.size() > 4096) andr_prof_flush(ctx);
}

void andr_prof_flush(andr_ctx *ctx) {
  ProfState &P = andr_prof(ctx);
  if (P.pending.empty()) return;
  (void)hipStreamSynchronize(ctx->stream);
  for (auto &r : P.pending) {
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, r.a, r.b);
    ProfAcc &A = P.acc[r.name];
    A.launches += 1;
    A.ms += ms;
    A.bytes += r.bytes;
    A.flops += r.flops;
    P.pool.push_back(r.a);
    P.pool.push_back(r.b);
  }
  P.pending.clear();
}

void andr_prof_adjust(andr_ctx *ctx, const char *name, double dbytes, double dflops) {
  ProfState &P = andr_prof(ctx);
  if (!P.on) return;
  ProfAcc &A = P.acc[name];
  A.bytes += dbytes;
  A.flops += dflops;
}

extern "C" int andr_profile(andr_ctx *ctx, int32_t enable) {
  if (!ctx) return ANDR_ERR_ARG;
  andr_prof_flush(ctx);
  ProfState &P = andr_prof(ctx);
  P.on = enable != 0;
  P.acc.clear();
  return ANDR_OK;
}

extern "C" int andr_profile_read(andr_ctx *ctx, int32_t idx, char *name, int32_t namelen, int64_t *launches,
                                 double *ms, double *bytes, double *flops) {
  if (!ctx) return ANDR_ERR_ARG;
  andr_prof_flush(ctx);
  ProfState &P = andr_prof(ctx);
  if (idx < 0 || idx >= (int)P.acc.size()) return 0;
  auto it = P.acc.begin();
  std::advance(it, idx);
  if (name && namelen > 0) {
    std::strncpy(name, it->first.c_str(), namelen - 1);
    name[namelen - 1] = 0;
  }
  if (launches) *launches = it->second.launches;
  if (ms) *ms = it->second.ms;
  if (bytes) *bytes = it->second.bytes;
  if (flops) *flops = it->second.flops;
  return 1;
}

extern "C" int andr_autotune(andr_ctx *ctx, int32_t mode) {
  if (!ctx) return ANDR_ERR_ARG;
  if (mode != 0 && mode != 1 && mode != 32 && mode != 64 && mode != 128)
    return andr_fail(ctx, ANDR_ERR_ARG, "andr_autotune: mode %d (0, 1, 32, 64 or 128)", mode);
  ctx->autotune = mode == 1;
  ctx->gemv_n_tiles = mode > 1 ? mode / 16 : 4;
  return ANDR_OK;
}

extern "C" int andr_autotune_read(andr_ctx *ctx, int32_t idx, int32_t *log2_rows, int32_t *log2_cols,
                                  int32_t *rows_per_wave, double *ms) {
  if (!ctx || idx < 0 || idx >= (int)ctx->tune_gemv_n.size()) return 0;
  auto it = ctx->tune_gemv_n.begin();
  std::advance(it, idx);
  if (log2_rows) *log2_rows = it->first.first;
  if (log2_cols) *log2_cols = it->first.second;
  if (rows_per_wave) *rows_per_wave = 16 * it->second.first;
  if (ms) *ms = it->second.second;
  return 1;
}

extern "C" int andr_autotune_read_krylov(andr_ctx *ctx, int32_t idx, int32_t *log2_row_waves, int32_t *target_waves,
                                         double *ms) {
  if (!ctx || idx < 0 || idx >= (int)ctx->tune_krylov.size()) return 0;
  auto it = ctx->tune_krylov.begin();
  std::advance(it, idx);
  if (log2_row_waves) *log2_row_waves = it->first;
  if (target_waves) *target_waves = it->second.first;
  if (ms) *ms = it->second.second;
  return 1;
}
