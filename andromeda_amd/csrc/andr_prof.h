// Optional per-kernel timing with HIP events on the context stream (off by default). Each timed
// launch carries its algorithmic bytes and flops so that achieved bandwidth / throughput can be
// reported against the roofline.
#pragma once
#include "andr_internal.h"

#include <map>
#include <string>
#include <vector>

struct ProfRec {
  std::string name;
  hipEvent_t a, b;
  double bytes, flops;
};
struct ProfAcc {
  int64_t launches = 0;
  double ms = 0.0, bytes = 0.0, flops = 0.0;
};
struct ProfState {
  bool on = false;
  std::vector<ProfRec> pending;
  std::vector<hipEvent_t> pool;
  std::map<std::string, ProfAcc> acc;
};
ProfState &andr_prof(andr_ctx *ctx);
// returns a token (index into pending) or -1 when profiling is off
int andr_prof_begin(andr_ctx *ctx, const char *name, double bytes, double flops);
void andr_prof_end(andr_ctx *ctx, int token);
void andr_prof_flush(andr_ctx *ctx);
// corrects the totals of a kernel name by dbytes and dflops (work the launches skipped on the
// device, known to the host only afterwards); no effect when profiling is off
void andr_prof_adjust(andr_ctx *ctx, const char *name, double dbytes, double dflops);
