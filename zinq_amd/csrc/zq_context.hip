// Context lifetime, error reporting, scratch memory, kernel timing, and ERI residency.
//
// The ERI upload path replaces what the reference does implicitly by keeping the dense n^4 tensor in
// host memory (src/molecular_system.zig:115-122 -> src/libint.cpp:161-167): here the unique integrals are
// re-laid into 8x8 AO-block quartets (see zq_internal.h) on the device and stay resident in HBM.

#include <algorithm>
#include <atomic>
#include <cstring>
#include <thread>

#include <shared_mutex>

#include "zq_internal.h"

int zq_build_jk_plan(zq_ctx *ctx);  // zq_jk.hip
int zq_jk_buffers(zq_ctx *ctx);     // zq_jk.hip

int zq_fail(zq_ctx *ctx, int code, const std::string &msg) {
  if (ctx) ctx->err = msg;
  return code;
}

int zq_check(zq_ctx *ctx, hipError_t e, const char *what) {
  if (e == hipSuccess) return ZQ_OK;
  const int code = (e == hipErrorOutOfMemory) ? ZQ_ERR_OOM : ZQ_ERR_LAUNCH;
  return zq_fail(ctx, code, std::string(what) + ": " + hipGetErrorString(e));
}

namespace {
std::shared_mutex g_gate;
thread_local int tl_depth = 0;     // shared holds of this thread (outermost one locks)
thread_local bool tl_excl = false;  // this thread is inside a capture window
}  // namespace

ZqGate::ZqGate() {
  if (!tl_excl && tl_depth++ == 0) g_gate.lock_shared();
}
ZqGate::~ZqGate() {
  if (!tl_excl && --tl_depth == 0) g_gate.unlock_shared();
}
ZqCaptureWindow::ZqCaptureWindow() : saved(tl_depth) {
  if (saved) {
    tl_depth = 0;
    g_gate.unlock_shared();
  }
  g_gate.lock();
  tl_excl = true;
}
ZqCaptureWindow::~ZqCaptureWindow() {
  tl_excl = false;
  g_gate.unlock();
  if (saved) {
    g_gate.lock_shared();
    tl_depth = saved;
  }
}

void *zq_scratch(zq_ctx *ctx, const std::string &name, size_t bytes, bool *fresh, bool captured) {
  DevBuf &b = ctx->scratch[name];
  if (fresh) *fresh = false;
  if (b.bytes >= bytes && b.ptr) return b.ptr;
  if (b.ptr) (void)hipFree(b.ptr);
  b.ptr = nullptr;
  b.bytes = 0;
  if (captured) ++ctx->scratch_gen;
  if (hipMalloc(&b.ptr, std::max<size_t>(bytes, 256)) != hipSuccess) {
    b.ptr = nullptr;
    (void)hipGetLastError();
    return nullptr;
  }
  b.bytes = bytes;
  if (fresh) *fresh = true;
  return b.ptr;
}

void *zq_pinned(zq_ctx *ctx, const std::string &name, size_t bytes) {
  DevBuf &b = ctx->pinned[name];
  if (b.bytes >= bytes && b.ptr) return b.ptr;
  if (b.ptr) (void)hipHostFree(b.ptr);
  b.ptr = nullptr;
  b.bytes = 0;
  if (hipHostMalloc(&b.ptr, std::max<size_t>(bytes, 256), hipHostMallocDefault) != hipSuccess) {
    b.ptr = nullptr;
    (void)hipGetLastError();
    return nullptr;
  }
  b.bytes = bytes;
  return b.ptr;
}

void zq_prof_begin(zq_ctx *ctx, const char *name) {
  if (!ctx->profile) return;
  zq_ctx::Ev ev;
  ev.name = name;
  (void)hipEventCreate(&ev.a);
  (void)hipEventCreate(&ev.b);
  (void)hipEventRecord(ev.a, ctx->stream);
  ctx->events.push_back(ev);
}

void zq_prof_end(zq_ctx *ctx) {
  if (!ctx->profile || ctx->events.empty()) return;
  (void)hipEventRecord(ctx->events.back().b, ctx->stream);
}

static void prof_collect(zq_ctx *ctx) {
  if (ctx->events.empty()) return;
  (void)hipStreamSynchronize(ctx->stream);
  for (auto &ev : ctx->events) {
    float ms = 0;
    if (hipEventElapsedTime(&ms, ev.a, ev.b) == hipSuccess) {
      auto &p = ctx->prof[ev.name];
      p.first += ms;
      p.second += 1;
    } else {
      (void)hipGetLastError();  // an unrecorded pair must not surface as the next launch's error
    }
    (void)hipEventDestroy(ev.a);
    (void)hipEventDestroy(ev.b);
  }
  ctx->events.clear();
}

void zq_free_eri(zq_ctx *ctx) {
  if (ctx->eri && ctx->eri_owned) (void)hipFree(ctx->eri);
  if (ctx->eri32 && ctx->eri32_owned) (void)hipFree(ctx->eri32);
  if (ctx->plan_owned)
    for (void *p : {(void *)ctx->d_items, (void *)ctx->d_csr_ptr, (void *)ctx->d_csr_idx})
      if (p) (void)hipFree(p);
  ctx->eri_owned = ctx->eri32_owned = ctx->plan_owned = true;
  ctx->set_id = 0;
  // the records, P, H and F buffers stay (grow-only); zq_destroy frees them
  ctx->eri = nullptr, ctx->eri32 = nullptr, ctx->d_items = nullptr;
  ctx->d_csr_ptr = nullptr, ctx->d_csr_idx = nullptr;
  ctx->n = ctx->npad = ctx->nb = ctx->Nb = 0;
  ctx->nquartets = 0;
  ctx->nitems = 0;
  ctx->nrecords = 0;
}

// One workgroup per block quartet; each element gathers its canonical packed source.
__global__ __launch_bounds__(256) void zq_pack_quartets(const double *__restrict__ canon, double *__restrict__ out,
                                                        int n, size_t nquartets) {
  const size_t qi = blockIdx.x;
  if (qi >= nquartets) return;
  // quartet index -> (Ib, k): Ib(Ib+1)/2 <= qi < (Ib+1)(Ib+2)/2
  size_t Ib = size_t((sqrt(8.0 * double(qi) + 1.0) - 1.0) * 0.5);
  while (Ib * (Ib + 1) / 2 > qi) --Ib;
  while ((Ib + 1) * (Ib + 2) / 2 <= qi) ++Ib;
  int k = int(qi - Ib * (Ib + 1) / 2);
  int bp, bq;
  zq_unpair(int(Ib), bp, bq);
  int bs = 0;
  for (;; ++bs) {
    const int len = zq_run_len(bp, bq, bs);
    if (k < len) break;
    k -= len;
  }
  const int br = bs + k;
  const int Jb = zq_bpair(br, bs);
  double *dst = out + qi * ZQ_QUARTET;
  for (int e = threadIdx.x; e < ZQ_QUARTET; e += blockDim.x) {
    const int rs = e >> 6, pq = e & 63;
    const int r = rs >> 3, s = rs & 7, p = pq >> 3, q = pq & 7;
    double v = 0.0;
    const bool dup = (bp == bq && p < q) || (br == bs && r < s) || (int(Ib) == Jb && pq < rs);
    const int P = bp * 8 + p, Q = bq * 8 + q, R = br * 8 + r, S = bs * 8 + s;
    if (!dup && P < n && Q < n && R < n && S < n) {
      const size_t I = size_t(P) * (P + 1) / 2 + Q, J = size_t(R) * (R + 1) / 2 + S;
      const size_t hi = I > J ? I : J, lo = I > J ? J : I;
      v = canon[hi * (hi + 1) / 2 + lo];
    }
    dst[e] = v;
  }
}

static int load_canonical(zq_ctx *ctx, const double *canon, size_t n) {
  if (!ctx || !canon || n == 0 || n > 4096) return zq_fail(ctx, ZQ_ERR_INVALID, "zq_eri_load: bad arguments");
  ZQ_TRY(ctx, hipSetDevice(ctx->device));
  zq_free_eri(ctx);
  ++ctx->eri_gen;
  ctx->n = int(n);
  ctx->nb = int((n + ZQ_BLK - 1) / ZQ_BLK);
  ctx->npad = ctx->nb * ZQ_BLK;
  ctx->Nb = ctx->nb * (ctx->nb + 1) / 2;
  ctx->nquartets = size_t(ctx->Nb) * (ctx->Nb + 1) / 2;
  const size_t np = n * (n + 1) / 2, nu = np * (np + 1) / 2;
  double *d_canon = nullptr;
  ZQ_TRY(ctx, hipMalloc(&d_canon, nu * sizeof(double)));
  hipError_t e = hipMemcpyAsync(d_canon, canon, nu * sizeof(double), hipMemcpyHostToDevice, ctx->stream);
  if (e == hipSuccess) e = hipMalloc(&ctx->eri, ctx->nquartets * ZQ_QUARTET * sizeof(double));
  if (e != hipSuccess) {
    (void)hipFree(d_canon);
    return zq_check(ctx, e, "zq_eri_load: upload");
  }
  zq_prof_begin(ctx, "pack_quartets");
  hipLaunchKernelGGL(zq_pack_quartets, dim3(unsigned(ctx->nquartets)), dim3(256), 0, ctx->stream, d_canon, ctx->eri,
                     int(n), ctx->nquartets);
  zq_prof_end(ctx);
  e = hipStreamSynchronize(ctx->stream);
  (void)hipFree(d_canon);
  if (e != hipSuccess) return zq_check(ctx, e, "zq_pack_quartets");
  ZQ_LAUNCH_CHECK(ctx, "zq_pack_quartets");
  return zq_build_jk_plan(ctx);
}

extern "C" {

zq_ctx *zq_create(int device) {
  ZqGate _gate;
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || device < 0 || device >= count) return nullptr;
  if (hipSetDevice(device) != hipSuccess) return nullptr;
  zq_noalloc_setup(device);
  auto *ctx = new zq_ctx;
  ctx->device = device;
  if (hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess) {
    delete ctx;
    return nullptr;
  }
  return ctx;
}

void zq_destroy(zq_ctx *ctx) {
  ZqGate _gate;
  if (!ctx) return;
  (void)hipSetDevice(ctx->device);
  (void)hipStreamSynchronize(ctx->stream);
  prof_collect(ctx);
  zq_release_solver(ctx);
  zq_free_eri(ctx);
  for (double *p : {ctx->d_records, ctx->d_P, ctx->d_H, ctx->d_F})
    if (p) (void)hipFree(p);
  for (auto &kv : ctx->scratch)
    if (kv.second.ptr) (void)hipFree(kv.second.ptr);
  for (auto &kv : ctx->pinned)
    if (kv.second.ptr) (void)hipHostFree(kv.second.ptr);
  (void)hipStreamDestroy(ctx->stream);
  delete ctx;
}

const char *zq_last_error(const zq_ctx *ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int zq_eri_load_packed(zq_ctx *ctx, const double *g_packed, size_t n) {
  ZqGate _gate;
  return load_canonical(ctx, g_packed, n);
}

int zq_eri_load(zq_ctx *ctx, const double *g, size_t n) {
  ZqGate _gate;
  if (!ctx || !g || n == 0 || n > 4096) return zq_fail(ctx, ZQ_ERR_INVALID, "zq_eri_load: bad arguments");
  // dense physicist <pr|qs> = (pq|rs) -> canonical packed chemist, in parallel over bra pairs
  const size_t np = n * (n + 1) / 2, nu = np * (np + 1) / 2;
  std::vector<double> canon;
  try {
    canon.resize(nu);
  } catch (...) {
    return zq_fail(ctx, ZQ_ERR_OOM, "zq_eri_load: host allocation");
  }
  const unsigned nth = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
  std::vector<std::thread> th;
  for (unsigned t = 0; t < nth; ++t)
    th.emplace_back([&, t]() {
      for (size_t p = t; p < n; p += nth)
        for (size_t q = 0; q <= p; ++q) {
          const size_t I = p * (p + 1) / 2 + q;
          for (size_t r = 0; r <= p; ++r)
            for (size_t s = 0; s <= r; ++s) {
              const size_t J = r * (r + 1) / 2 + s;
              if (J > I) break;
              canon[I * (I + 1) / 2 + J] = g[((p * n + r) * n + q) * n + s];
            }
        }
    });
  for (auto &x : th) x.join();
  return load_canonical(ctx, canon.data(), n);
}

void zq_profile_enable(zq_ctx *ctx, int on) {
  ZqGate _gate;
  if (ctx) ctx->profile = on != 0;
}

void zq_profile_reset(zq_ctx *ctx) {
  ZqGate _gate;
  if (!ctx) return;
  prof_collect(ctx);
  ctx->prof.clear();
}

int zq_profile_get(zq_ctx *ctx, const char *prefix, double *ms, int *launches) {
  ZqGate _gate;
  if (!ctx || !prefix) return ZQ_ERR_INVALID;
  prof_collect(ctx);
  double t = 0;
  int c = 0;
  const size_t lp = std::strlen(prefix);
  for (auto &kv : ctx->prof)
    if (kv.first.compare(0, lp, prefix) == 0) t += kv.second.first, c += kv.second.second;
  if (ms) *ms = t;
  if (launches) *launches = c;
  return ZQ_OK;
}

zq_eri_set *zq_eri_set_create(int device, const double *g_packed, size_t n, int with_fp32) {
  ZqGate _gate;
  int count = 0;
  if (!g_packed || n == 0 || n > 4096 || hipGetDeviceCount(&count) != hipSuccess || device < 0 || device >= count)
    return nullptr;
  if (hipSetDevice(device) != hipSuccess) return nullptr;
  static std::atomic<unsigned long long> next_id{1};
  auto *s = new zq_eri_set;
  s->id = next_id.fetch_add(1);
  s->device = device;
  s->n = int(n);
  s->nb = int((n + ZQ_BLK - 1) / ZQ_BLK);
  s->npad = s->nb * ZQ_BLK;
  s->Nb = s->nb * (s->nb + 1) / 2;
  s->nquartets = size_t(s->Nb) * (s->Nb + 1) / 2;
  const size_t np = n * (n + 1) / 2, nu = np * (np + 1) / 2, cnt = s->nquartets * ZQ_QUARTET;
  double *d_canon = nullptr;
  hipStream_t st = nullptr;
  bool ok = hipStreamCreateWithFlags(&st, hipStreamNonBlocking) == hipSuccess &&
            hipMalloc(&d_canon, nu * sizeof(double)) == hipSuccess && hipMalloc(&s->eri, cnt * sizeof(double)) == hipSuccess &&
            hipMemcpyAsync(d_canon, g_packed, nu * sizeof(double), hipMemcpyHostToDevice, st) == hipSuccess;
  if (ok) {
    hipLaunchKernelGGL(zq_pack_quartets, dim3(unsigned(s->nquartets)), dim3(256), 0, st, d_canon, s->eri, int(n),
                       s->nquartets);
    ok = hipGetLastError() == hipSuccess;
  }
  if (ok && with_fp32) {
    ok = hipMalloc(&s->eri32, cnt * sizeof(float)) == hipSuccess;
    if (ok) {
      zq_narrow_eri(st, s->eri, cnt, s->eri32);
      ok = hipGetLastError() == hipSuccess;
    }
  }
  // the J/K plan of this size, once for every context that attaches
  if (ok) ok = zq_jk_plan_upload(s->Nb, s->nquartets, st, s->plan, s->err) == ZQ_OK;
  if (st) ok = (hipStreamSynchronize(st) == hipSuccess) && ok;
  if (d_canon) (void)hipFree(d_canon);
  if (st) (void)hipStreamDestroy(st);
  if (!ok) {
    (void)hipGetLastError();
    zq_eri_set_destroy(s);
    return nullptr;
  }
  return s;
}

int zq_mem_info(int device, size_t *free_bytes, size_t *total_bytes) {
  ZqGate _gate;
  if (!free_bytes || !total_bytes) return ZQ_ERR_INVALID;
  if (hipSetDevice(device) != hipSuccess) return ZQ_ERR_INVALID;
  return hipMemGetInfo(free_bytes, total_bytes) == hipSuccess ? ZQ_OK : ZQ_ERR_LAUNCH;
}

void zq_eri_set_destroy(zq_eri_set *s) {
  ZqGate _gate;
  if (!s) return;
  (void)hipSetDevice(s->device);
  if (s->eri) (void)hipFree(s->eri);
  if (s->eri32) (void)hipFree(s->eri32);
  zq_jk_plan_free(s->plan);
  delete s;
}

size_t zq_eri_set_bytes(const zq_eri_set *s) {
  if (!s) return 0;
  return s->nquartets * ZQ_QUARTET * (sizeof(double) + (s->eri32 ? sizeof(float) : 0));
}

int zq_attach(zq_ctx *ctx, zq_eri_set *s) {
  ZqGate _gate;
  if (!ctx || !s) return zq_fail(ctx, ZQ_ERR_INVALID, "zq_attach: null argument");
  if (s->device != ctx->device) return zq_fail(ctx, ZQ_ERR_INVALID, "zq_attach: set and context on different devices");
  ZQ_TRY(ctx, hipSetDevice(ctx->device));
  ZQ_TRY(ctx, hipStreamSynchronize(ctx->stream));
  if (ctx->set_id == s->id) return ZQ_OK;  // already attached: plan and buffers stay
  zq_free_eri(ctx);
  ++ctx->eri_gen;
  ctx->n = s->n, ctx->nb = s->nb, ctx->npad = s->npad, ctx->Nb = s->Nb, ctx->nquartets = s->nquartets;
  ctx->eri = s->eri, ctx->eri_owned = false;
  ctx->eri32 = s->eri32, ctx->eri32_owned = s->eri32 == nullptr;
  ctx->d_items = s->plan.items, ctx->d_csr_ptr = s->plan.ptr, ctx->d_csr_idx = s->plan.idx;
  ctx->nitems = s->plan.nitems, ctx->nrecords = s->plan.nrecords, ctx->plan_owned = false;
  if (int rc = zq_jk_buffers(ctx)) {
    zq_free_eri(ctx);
    return rc;
  }
  ctx->set_id = s->id;
  return ZQ_OK;
}

size_t zq_device_bytes(const zq_ctx *ctx) {
  if (!ctx) return 0;
  size_t b = ctx->eri_owned ? ctx->nquartets * ZQ_QUARTET * sizeof(double) : 0;
  b += (ctx->cap_records + ctx->cap_P + ctx->cap_H + ctx->cap_F) * sizeof(double);
  if (ctx->eri32 && ctx->eri32_owned) b += ctx->nquartets * ZQ_QUARTET * sizeof(float);
  for (auto &kv : ctx->scratch) b += kv.second.bytes;
  return b;
}

}  // extern "C"
