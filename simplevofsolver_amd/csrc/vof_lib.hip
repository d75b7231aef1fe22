// vof_lib.hip -- host side of the gfx950 SimpleVOFSolver time step: the
// reference's entry points (include/vof_solver.h) over device-resident
// arrays, the NPY codec, and the launch sequences of the kernels in
// vof_kernels.h.  Built once per NDIMS into libvofhip{2,3}d.so.
#include <hip/hip_runtime.h>

#include <math.h>
#include <float.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <errno.h>
#include <sys/stat.h>

#include <atomic>
#include <chrono>
#include <condition_variable>
#include <functional>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <utility>
#include <vector>

#include "vof_solver.h"
#include "vof_kernels.h"
#include "vof_adjoint.h"

using vofk::Grid;
using vofk::FftPlan;

// ---------------------------------------------------------------------------
// errors

static thread_local std::string g_err;

static int fail(const char * fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_err = buf;
  return 1;
}

#define HIPCHK(call)                                                               \
  do {                                                                             \
    hipError_t e_ = (call);                                                        \
    if (e_ != hipSuccess) return fail("%s:%d %s: %s", __FILE__, __LINE__, #call,   \
                                      hipGetErrorString(e_));                      \
  } while (0)

// release paths: a failure to free is not reported (the handle is gone either way)
static void dev_free(void * p) {
  if (p) (void)hipFree(p);
}
static void host_free(void * p) {
  if (p) (void)hipHostFree(p);
}

#define LAUNCHCHK()                                                                \
  do {                                                                             \
    hipError_t e_ = hipGetLastError();                                             \
    if (e_ != hipSuccess) return fail("%s:%d kernel launch: %s", __FILE__, __LINE__, \
                                      hipGetErrorString(e_));                      \
  } while (0)

extern "C" const char * vof_last_error(void) { return g_err.c_str(); }

// ---------------------------------------------------------------------------
// NPY v1 codec (the format checks of src/fileio.c:185-245: dtype, shape,
// C order)

namespace npy {

static std::string path_of(const char * dir, const char * name) {
  std::string p(dir);
  if (!p.empty() && p.back() != '/') p += '/';
  p += name;
  p += ".npy";
  return p;
}

// returns 0 and fills data (nbytes) when the header matches exactly
static int read(const char * dir, const char * name, const char * dtype,
                const std::vector<size_t> & shape, void * data, size_t elsize) {
  const std::string p = path_of(dir, name);
  FILE * fp = fopen(p.c_str(), "rb");
  if (!fp) return fail("%s: %s", p.c_str(), strerror(errno));
  unsigned char pre[10];
  if (fread(pre, 1, 10, fp) != 10 || memcmp(pre, "\x93NUMPY", 6) != 0) {
    fclose(fp);
    return fail("%s: not an NPY file", p.c_str());
  }
  const int major = pre[6];
  size_t hlen = 0, pre_len = 10;
  if (major == 1) {
    hlen = (size_t)pre[8] | ((size_t)pre[9] << 8);
  } else if (major == 2 || major == 3) {
    unsigned char b[2];
    if (fread(b, 1, 2, fp) != 2) { fclose(fp); return fail("%s: short header", p.c_str()); }
    hlen = (size_t)pre[8] | ((size_t)pre[9] << 8) | ((size_t)b[0] << 16) | ((size_t)b[1] << 24);
    pre_len = 12;
  } else {
    fclose(fp);
    return fail("%s: unsupported NPY version %d", p.c_str(), major);
  }
  std::string hdr(hlen, '\0');
  if (fread(&hdr[0], 1, hlen, fp) != hlen) { fclose(fp); return fail("%s: short header", p.c_str()); }
  // descr
  const size_t d = hdr.find("'descr'");
  const size_t f = hdr.find("'fortran_order'");
  const size_t s = hdr.find("'shape'");
  if (d == std::string::npos || f == std::string::npos || s == std::string::npos) {
    fclose(fp);
    return fail("%s: malformed header", p.c_str());
  }
  {
    const size_t q0 = hdr.find('\'', d + 7);
    const size_t q1 = hdr.find('\'', q0 + 1);
    const std::string dt = hdr.substr(q0, q1 - q0 + 1);
    if (dt != dtype) {
      fclose(fp);
      return fail("%s: dtype %s expected, %s obtained", p.c_str(), dtype, dt.c_str());
    }
  }
  {
    const size_t c = hdr.find(':', f);
    size_t v = hdr.find_first_not_of(" ", c + 1);
    if (hdr.compare(v, 5, "False") != 0) {
      fclose(fp);
      return fail("%s: Fortran order is not supported", p.c_str());
    }
  }
  {
    const size_t o = hdr.find('(', s);
    const size_t e = hdr.find(')', o);
    std::vector<size_t> got;
    size_t pos = o + 1;
    while (pos < e) {
      while (pos < e && (hdr[pos] == ' ' || hdr[pos] == ',')) pos++;
      if (pos >= e) break;
      got.push_back((size_t)strtoull(hdr.c_str() + pos, nullptr, 10));
      while (pos < e && hdr[pos] != ',') pos++;
    }
    if (got != shape) {
      std::string a, b;
      for (size_t x : shape) a += std::to_string(x) + ",";
      for (size_t x : got) b += std::to_string(x) + ",";
      fclose(fp);
      return fail("%s: shape (%s) expected, (%s) obtained", p.c_str(), a.c_str(), b.c_str());
    }
  }
  (void)pre_len;
  size_t n = 1;
  for (size_t x : shape) n *= x;
  if (fread(data, elsize, n, fp) != n) {
    fclose(fp);
    return fail("%s: data truncated", p.c_str());
  }
  fclose(fp);
  return 0;
}

static int write(const char * dir, const char * name, const char * dtype,
                 const std::vector<size_t> & shape, const void * data, size_t elsize) {
  const std::string p = path_of(dir, name);
  std::string sh = "(";
  for (size_t m = 0; m < shape.size(); m++) {
    sh += std::to_string(shape[m]);
    if (shape.size() == 1 || m + 1 < shape.size()) sh += ",";
    if (m + 1 < shape.size()) sh += " ";
  }
  sh += ")";
  std::string hdr = std::string("{'descr': ") + dtype + ", 'fortran_order': False, 'shape': " + sh + ", }";
  // pad with spaces so that magic + length + header + '\n' is a multiple of 64
  const size_t base = 10 + hdr.size() + 1;
  hdr.append((64 - base % 64) % 64, ' ');
  hdr += '\n';
  FILE * fp = fopen(p.c_str(), "wb");
  if (!fp) return fail("%s: %s", p.c_str(), strerror(errno));
  unsigned char pre[10] = {0x93, 'N', 'U', 'M', 'P', 'Y', 1, 0, 0, 0};
  pre[8] = (unsigned char)(hdr.size() & 0xff);
  pre[9] = (unsigned char)((hdr.size() >> 8) & 0xff);
  size_t n = 1;
  for (size_t x : shape) n *= x;
  const bool ok = fwrite(pre, 1, 10, fp) == 10 && fwrite(hdr.data(), 1, hdr.size(), fp) == hdr.size() &&
                  fwrite(data, elsize, n, fp) == n;
  fclose(fp);
  if (!ok) return fail("%s: write failed", p.c_str());
  return 0;
}

}  // namespace npy

static const char kNpyDouble[] = "'<f8'";
static const char kNpySizeT[] = "'<u8'";

// ---------------------------------------------------------------------------
// solver context (lives behind domain_t.info)

enum {
  P_RA, P_PR, P_WE, P_COEF_ADV, P_COEF_DIF, P_BUOY, P_UY_XM, P_UY_XP, P_UZ_XM, P_UZ_XP,
  P_T_XM, P_T_XP, P_COUNT
};
static const char * const kParamNames[P_COUNT] = {
  "Ra", "Pr", "We", "coef_dt_adv", "coef_dt_dif", "add_buoyancy", "uy_xm", "uy_xp",
  "uz_xm", "uz_xp", "t_xm", "t_xp",
};

struct Recorder;

struct vof_ctx {
  Recorder * rec;              // non-null: launches are recorded for a batch, not issued
  bool tab_shared;             // Thomas tables live in a batch's block
  double ** tape;              // non-null: stage states (vof, ux, uy) are copied here per RK stage
  double adj_floor;            // |grad vof| dx below which corner normals are constants in the adjoint
  int nx, ny, nz;              // nz = 1 in 2D
  double dy, dz;
  bool x_uniform;
  hipStream_t stream;
  double prm[P_COUNT];
  bool prm_set[P_COUNT];
  int schedule;
  int fuse;                              // F_* stages fused in schedule 1
  int fft_elems;                         // LDS complex elements per FFT buffer (1024 / 2048)
  int lines_lpb;                         // lines per work-group of the x solver (0 = auto)
  int rhs_rows;                          // rows marched per thread by the source kernel (1 = one DOF per thread)
  int iface_planes;                      // z planes marched per work-group by the 3D interface kernel (0 = blocks)
  int fft_radix;                         // 16: 256-point transforms on the register radix-16 path; 4: mixed radix
  int thomas_rows;                       // rows per prefetched chunk of the spectral Thomas kernel (8, 16, 32)
  int thomas_part;                       // 1: partitioned spectral x solve (k_thomas_part) where 4 <= nx <= 256
  int iface_overlap;                     // 1: the next stage's interface kernel runs on `side` (see integrate)
  hipStream_t side;                      // second stream of the fused schedule (default priority)
  hipEvent_t ev_rhs, ev_iface, ev_main;  // side launch point (main); interface kernel done (side); main so far
  const interface_t * pending_iface;     // the next stage's interface kernel, to go on the side stream
  bool iface_ready;                      // the first stage's interface kernel is already on the main stream
  double * d_vof_spare;                  // second VOF buffer of the fused update
  size_t vof_bytes;
  double * d_fspare[4];                  // second buffer of ux, uy, uz, t (one-kernel predictor)
  // decision record / replay of the 2D step (vof_set_freeze): one slot per
  // RK stage, each one device block {corner, cell, x face, y face flags,
  // corner normals}
  int freeze_mode;                       // 0 off, 1 record, 2 replay
  size_t freeze_pos, freeze_n;           // next stage slot, slots recorded
  std::vector<void *> freeze_slots;
  int pred_rows;                         // rows marched per thread by the one-kernel predictor (1, 2, 4, 8)
  Grid grid;
  double * d_dxf;
  double * d_dxc;
  std::vector<double> lapx_ux, lapx_c;   // 3 per point, predict/{ux,uy,t}.c init_lap
  double * d_lapx_ux;
  double * d_lapx_c;
  // predictor Thomas tables: one slot per (field, rkstep)
  size_t tab_stride;
  size_t tab_rk;               // doubles between the tables of consecutive RK stages
  double * d_tab;
  double * h_tab;                        // pinned
  hipEvent_t tab_done[4 * 3];
  // Poisson
  int nh;
  FftPlan plan_y, plan_z;
  double2 * d_twy;
  double2 * d_twz;
  double * d_pl;                         // x operator lower / upper (dft.c:160-190)
  double * d_pu;
  double * d_ey;                         // eigenvalues in y, z (dft.c:263-291)
  double * d_ez;
  double * d_stab;                       // factors of the (0, 0) line: c0, rlast, flag, val[nx], v[nx]
  double2 * d_spec;
  // mixed-precision Poisson (refinements < 0: plain FP64)
  int poisson_refine;
  float2 * d_spec32;
  float2 * d_twy32;
  float2 * d_twz32;
  double * d_vbuf;
  bool vbuf_ready;             // the Thomas back-substitution factors are in d_vbuf
  // scratch
  double * d_delta;                      // x1 pencil of increments
  double * d_part;
  int npart;
  double * h_scalar;                     // pinned
  // statistics accumulators (statistics.c), allocated by statistics.init
  double * d_stat[9];
  size_t stat_num;
  double stat_rate, stat_next, stat_m_dif, stat_t_dif;
  // per-kernel timing with HIP events on this context's stream
  bool prof_on;
  std::vector<hipEvent_t> ev_pool;
  size_t ev_next;
  struct Pending { const char * name; hipEvent_t a, b; };
  std::vector<Pending> pending;
  std::vector<std::string> prof_names;
  std::vector<double> prof_ms;
  std::vector<long> prof_cnt;
};

static vof_ctx * ctx_of(const domain_t * domain) {
  return domain ? (vof_ctx *)domain->info : nullptr;
}

#define CTX_OR_FAIL(dom)                                                           \
  vof_ctx * ctx = ctx_of(dom);                                                     \
  if (!ctx) return fail("%s: domain is not initialised", __func__)

static int param_get(vof_ctx * ctx, int id, double * out) {
  if (ctx->prm_set[id]) {
    *out = ctx->prm[id];
    return 0;
  }
  // config.get_double (src/config.c:14-42)
  const char * s = getenv(kParamNames[id]);
  if (!s) return fail("%s not found (environment variable or vof_set_param)", kParamNames[id]);
  errno = 0;
  const double v = strtod(s, nullptr);
  if (errno != 0) return fail("%s: invalid value as double", kParamNames[id]);
  *out = v;
  return 0;
}

// ---------------------------------------------------------------------------
// launch helpers

static dim3 grid3(int ni, int nj, int nk, dim3 blk) {
  return dim3((unsigned)((ni + blk.x - 1) / blk.x), (unsigned)((nj + blk.y - 1) / blk.y),
              (unsigned)(NDIMS == 3 ? nk : 1));
}

static const dim3 kBlk(64, 4, 1);


static int prof_flush(vof_ctx * ctx) {
  if (ctx->pending.empty()) return 0;
  HIPCHK(hipStreamSynchronize(ctx->stream));
  for (const auto & p : ctx->pending) {
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, p.a, p.b));
    size_t m = 0;
    while (m < ctx->prof_names.size() && ctx->prof_names[m] != p.name) m++;
    if (m == ctx->prof_names.size()) {
      ctx->prof_names.push_back(p.name);
      ctx->prof_ms.push_back(0.);
      ctx->prof_cnt.push_back(0);
    }
    ctx->prof_ms[m] += ms;
    ctx->prof_cnt[m] += 1;
  }
  ctx->pending.clear();
  ctx->ev_next = 0;
  return 0;
}

static hipEvent_t prof_begin(vof_ctx * ctx) {
  if (!ctx->prof_on) return nullptr;
  if (ctx->ev_next + 2 > ctx->ev_pool.size() && prof_flush(ctx)) return nullptr;
  hipEvent_t e = ctx->ev_pool[ctx->ev_next++];
  if (hipEventRecord(e, ctx->stream) != hipSuccess) return nullptr;
  return e;
}

static int prof_end(vof_ctx * ctx, const char * name, hipEvent_t e0) {
  hipEvent_t e1 = ctx->ev_pool[ctx->ev_next++];
  HIPCHK(hipEventRecord(e1, ctx->stream));
  ctx->pending.push_back({name, e0, e1});
  return 0;
}

// ---------------------------------------------------------------------------
// Batched execution.  A batch advances many independent cases (each a full
// context of its own, any sizes) with one launch per kernel of the step.  The
// host code of the step runs once per case in record mode: every launch is
// captured with its grid and packed arguments instead of being issued.  The
// i-th launches of all cases (the same kernel instantiation) then become one
// launch of k_batched, whose work-groups are the concatenation of the cases'
// work-groups; each work-group finds its case by binary search over the
// prefix sums and runs the kernel body with that case's arguments and its
// own tile coordinates.  Results are those of the single-case launches, bit
// for bit: every tile executes the same body on the same data.

template <class... A> struct Pack;
template <> struct Pack<> {};
template <class H, class... T> struct Pack<H, T...> {
  H h;
  Pack<T...> t;
};

template <size_t I, class H, class... T>
__host__ __device__ constexpr const auto & pack_get(const Pack<H, T...> & p) {
  if constexpr (I == 0)
    return p.h;
  else
    return pack_get<I - 1>(p.t);
}

template <class H, class... T>
static Pack<H, T...> make_pack(H h, T... t) {
  Pack<H, T...> p;
  p.h = h;
  if constexpr (sizeof...(T) > 0) p.t = make_pack(t...);
  return p;
}

template <class... A> struct BatchEntry {
  unsigned first;              // first work-group of this case in the batched launch
  unsigned gx, gy, gz;         // the case's own grid
  Pack<A...> args;
};

template <auto Body, class... A, size_t... I>
__device__ __forceinline__ void apply_body(const vofk::VBlock & vb, const Pack<A...> & p,
                                           std::index_sequence<I...>) {
  Body(vb, pack_get<I>(p)...);
}

template <auto Body, unsigned BOUNDS, class... A>
__global__ __launch_bounds__(BOUNDS) void k_batched(const BatchEntry<A...> * __restrict__ e, int ncases) {
  int lo = 0, hi = ncases - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (e[mid].first <= blockIdx.x)
      lo = mid;
    else
      hi = mid - 1;
  }
  const BatchEntry<A...> & E = e[lo];
  const unsigned l = blockIdx.x - E.first;
  vofk::VBlock vb;
  vb.gx = E.gx;
  vb.gy = E.gy;
  vb.gz = E.gz;
  vb.x = l % E.gx;
  vb.y = (l / E.gx) % E.gy;
  vb.z = l / (E.gx * E.gy);
  apply_body<Body>(vb, E.args, std::index_sequence_for<A...>{});
}

struct vof_batch;

// one recorded launch; `merge` is specific to the kernel instantiation and
// doubles as its identity
struct LaunchRec {
  size_t (*merge)(vof_batch * B, size_t idx, const std::vector<int> & cases, unsigned char * host,
                  const unsigned char * dev);
  const char * name;
  dim3 grid, block;
  size_t shm;
  std::vector<unsigned char> pack;
};

struct Recorder {
  std::vector<LaunchRec> recs;
};

// Host worker threads of a batch: run(n, f) calls f(0 .. n-1) spread over
// the workers and the calling thread and returns when all calls are done.
// Recording a case's launches touches only that case's context, recorder
// and table slots, so the cases of a phase are recorded in parallel.
struct HostPool {
  std::vector<std::thread> th;
  std::mutex m;
  std::condition_variable cv, done;
  const std::function<void(int)> * job = nullptr;
  int njobs = 0;
  std::atomic<int> next{0};
  int active = 0;
  unsigned gen = 0;
  bool stop = false;

  void work() {
    for (int i; (i = next.fetch_add(1)) < njobs;) (*job)(i);
  }
  void loop() {
    unsigned seen = 0;
    std::unique_lock<std::mutex> l(m);
    for (;;) {
      cv.wait(l, [&] { return stop || gen != seen; });
      if (stop) return;
      seen = gen;
      l.unlock();
      work();
      l.lock();
      if (--active == 0) done.notify_all();
    }
  }
  void start(int nthreads) {
    for (int t = 0; t < nthreads; t++) th.emplace_back([this] { loop(); });
  }
  void run(int n, const std::function<void(int)> & f) {
    if (th.empty() || n < 2) {
      for (int i = 0; i < n; i++) f(i);
      return;
    }
    {
      std::lock_guard<std::mutex> l(m);
      job = &f;
      njobs = n;
      next = 0;
      active = (int)th.size();
      gen++;
    }
    cv.notify_all();
    work();
    std::unique_lock<std::mutex> l(m);
    done.wait(l, [&] { return active == 0; });
    job = nullptr;
  }
  ~HostPool() {
    {
      std::lock_guard<std::mutex> l(m);
      stop = true;
    }
    cv.notify_all();
    for (auto & t : th) t.join();
  }
};

struct vof_batch {
  HostPool pool;
  std::vector<vof_ctx *> ctx;
  std::vector<const domain_t *> domain;
  std::vector<fluid_t *> fluid;
  std::vector<interface_t *> interface;
  std::vector<Recorder> rec;
  hipStream_t stream;
  // staging of the merged launches' argument tables, one pinned + device
  // pair per phase of the step (0: dt, 1..3: RK stages)
  unsigned char * h_stage[4];
  unsigned char * d_stage[4];
  size_t stage_bytes[4];
  hipEvent_t stage_done[4];
  // hipGraph of each phase (its copies and merged launches), captured once
  // and replayed while the phase's launch layout is unchanged
  bool graphs;
  hipGraphExec_t gexec[4];
  std::vector<size_t> gsig[4];
  size_t graph_captures, graph_launches;
  // Thomas tables of every case, contiguous (cases re-pointed into these)
  double * h_tab;
  double * d_tab;
  size_t tab_doubles;
  // dt gather
  double * d_dt;
  double * h_dt;
  const double ** d_dt_src;
  // a batch of 8 cases or more is 2 batches of alternate cases (4 from 16
  // cases on: case c in part c % 4), each with its own stream, tables and
  // graphs, stepped in turn: while the host reads one part's dt and records
  // its first stage, the device runs the other parts' stages.
  // part_cases[p][i] is the case number of case i of part p.  A batch with
  // parts holds no stream itself.
  std::vector<vof_batch *> parts;
  std::vector<std::vector<int>> part_cases;
};
static int batch_parts(int ncases) { return ncases >= 16 ? 4 : ncases >= 8 ? 2 : 1; }

// build (dev == nullptr: size only / fill host) or launch (dev != nullptr)
// the merged launch number idx of the listed cases (which all recorded the
// same kernel there); returns the bytes of its argument table
template <auto Body, unsigned BOUNDS, class... A>
static size_t merge_launch(vof_batch * B, size_t idx, const std::vector<int> & cases, unsigned char * host,
                           const unsigned char * dev) {
  using E = BatchEntry<A...>;
  const size_t n = cases.size();
  const size_t bytes = sizeof(E) * n;
  if (!dev) {
    if (host) {
      unsigned first = 0;
      for (size_t m = 0; m < n; m++) {
        const LaunchRec & r = B->rec[cases[m]].recs[idx];
        E e;
        e.first = first;
        e.gx = r.grid.x;
        e.gy = r.grid.y;
        e.gz = r.grid.z;
        memcpy(&e.args, r.pack.data(), sizeof(Pack<A...>));
        memcpy(host + sizeof(E) * m, &e, sizeof(E));
        first += r.grid.x * r.grid.y * r.grid.z;
      }
    }
    return bytes;
  }
  unsigned total = 0;
  size_t shm = 0;
  for (size_t m = 0; m < n; m++) {
    const LaunchRec & r = B->rec[cases[m]].recs[idx];
    total += r.grid.x * r.grid.y * r.grid.z;
    shm = r.shm > shm ? r.shm : shm;
  }
  const dim3 blk = B->rec[cases[0]].recs[idx].block;
  hipLaunchKernelGGL((k_batched<Body, BOUNDS, A...>), dim3(total), blk, shm, B->stream,
                     reinterpret_cast<const E *>(dev), (int)n);
  return hipGetLastError() == hipSuccess ? bytes : 0;
}

template <auto Body, unsigned BOUNDS, class... A>
static void record_launch(Recorder * rec, const char * name, dim3 g, dim3 b, size_t shm, A... args) {
  LaunchRec r;
  r.merge = &merge_launch<Body, BOUNDS, A...>;
  r.name = name;
  r.grid = g;
  r.block = b;
  r.shm = shm;
  const Pack<A...> p = make_pack(args...);
  r.pack.resize(sizeof(p));
  memcpy(r.pack.data(), &p, sizeof(p));
  rec->recs.push_back(std::move(r));
}

// launch of a batchable kernel: BODY is its device body, BOUNDS its
// work-group size bound; recorded instead of issued in record mode
#define KLB(ctx, name, BOUNDS, BODY, kern, g, b, shm, ...)                           \
  do {                                                                             \
    if ((ctx)->rec) {                                                              \
      record_launch<BODY, BOUNDS>((ctx)->rec, name, g, b, shm, __VA_ARGS__);        \
    } else {                                                                       \
      hipEvent_t e0_ = prof_begin(ctx);                                            \
      hipLaunchKernelGGL(kern, g, b, shm, (ctx)->stream, __VA_ARGS__);             \
      LAUNCHCHK();                                                                 \
      if (e0_ && prof_end(ctx, name, e0_)) return 1;                               \
    }                                                                              \
  } while (0)

// launch on the context stream, bracketed by events when profiling is on
#define KLN(ctx, name, kern, g, b, ...)                                            \
  do {                                                                             \
    if ((ctx)->rec) return fail("%s cannot run inside a batched step", name);      \
    hipEvent_t e0_ = prof_begin(ctx);                                              \
    hipLaunchKernelGGL(kern, g, b, 0, (ctx)->stream, __VA_ARGS__);                 \
    LAUNCHCHK();                                                                   \
    if (e0_ && prof_end(ctx, name, e0_)) return 1;                                 \
  } while (0)

// ---------------------------------------------------------------------------
// boundary conditions: halo copy, then x walls (src/halo.c:18-76 followed by
// assign_boundary_conditions_in_x of each fluid/boundary/*.c)

enum { W_DIRICHLET = 0, W_NEUMANN = 1 };

static int bc_apply(vof_ctx * ctx, double * a, int n0, int h, int walls_all_rows, int mode,
                    double vlo, double vhi, int xlo, int xhi) {
  const int ny = ctx->ny, nz = ctx->nz;
  const int n1 = ny + 2 * h;
#if NDIMS == 2
  const size_t cnt = (size_t)n0 * 2 * h;
#else
  const size_t cnt = (size_t)n0 * 2 * h * (nz + 2 * h) + (size_t)n0 * ny * 2 * h;
#endif
  {
    const unsigned nb = (unsigned)((cnt + 255) / 256);
    KLN(ctx, "k_halo", vofk::k_halo, dim3(nb), dim3(256), a, n0, ny, nz, h);
    LAUNCHCHK();
  }
  int r0, nr, p0, np;
  if (walls_all_rows) {
    r0 = 0; nr = n1;
#if NDIMS == 2
    p0 = 0; np = 1;
#else
    p0 = 0; np = nz + 2 * h;
#endif
  } else {
    r0 = h; nr = ny;
#if NDIMS == 2
    p0 = 0; np = 1;
#else
    p0 = h; np = nz;
#endif
  }
  const int tot = nr * np;
  KLN(ctx, "k_walls", vofk::k_walls, dim3((tot + 255) / 256), dim3(256), a, n0,
                     NDIMS == 3 ? n1 : 0, r0, nr, p0, np, xlo, xhi, mode, vlo, vhi);
  LAUNCHCHK();
  return 0;
}

static int bc_ux(vof_ctx * ctx, double * ux) {
  return bc_apply(ctx, ux, ctx->nx + 1, 1, 0, W_DIRICHLET, 0., 0., 0, ctx->nx);
}
static int bc_uy(vof_ctx * ctx, double * uy) {
  return bc_apply(ctx, uy, ctx->nx + 2, 1, 0, W_DIRICHLET, ctx->prm[P_UY_XM], ctx->prm[P_UY_XP],
                  0, ctx->nx + 1);
}
static int bc_uz(vof_ctx * ctx, double * uz) {
  return bc_apply(ctx, uz, ctx->nx + 2, 1, 0, W_DIRICHLET, ctx->prm[P_UZ_XM], ctx->prm[P_UZ_XP],
                  0, ctx->nx + 1);
}
static int bc_neumann(vof_ctx * ctx, double * p) {
  return bc_apply(ctx, p, ctx->nx + 2, 1, 0, W_NEUMANN, 0., 0., 0, ctx->nx + 1);
}
static int bc_t(vof_ctx * ctx, double * t) {
  return bc_apply(ctx, t, ctx->nx + 2, 1, 0, W_DIRICHLET, ctx->prm[P_T_XM], ctx->prm[P_T_XP], 0,
                  ctx->nx + 1);
}
static int bc_vof(vof_ctx * ctx, double * vof) {
  return bc_apply(ctx, vof, ctx->nx + 2, 2, 1, W_DIRICHLET, 0., 0., 0, ctx->nx + 1);
}

// ---------------------------------------------------------------------------
// arrays (src/array.c:10-46): padded sizes from the nadds tables

struct Nadds { int v[3][2]; };

static Nadds nadds_of(const char * name) {
  // tools/define_arrays.py:201-229
  struct E { const char * n; int a[3][2]; };
  static const E tab[] = {
    {"ux", {{0, 1}, {1, 1}, {1, 1}}},     {"uy", {{1, 1}, {1, 1}, {1, 1}}},
    {"uz", {{1, 1}, {1, 1}, {1, 1}}},     {"p", {{1, 1}, {1, 1}, {1, 1}}},
    {"t", {{1, 1}, {1, 1}, {1, 1}}},      {"psi", {{1, 1}, {1, 1}, {1, 1}}},
    {"srcux", {{-1, 0}, {0, 0}, {0, 0}}}, {"srcuy", {{0, 0}, {0, 0}, {0, 0}}},
    {"srcuz", {{0, 0}, {0, 0}, {0, 0}}},  {"srct", {{0, 0}, {0, 0}, {0, 0}}},
    {"ifrcx", {{-1, 0}, {0, 0}, {0, 0}}}, {"ifrcy", {{0, 0}, {0, 0}, {0, 0}}},
    {"ifrcz", {{0, 0}, {0, 0}, {0, 0}}},  {"vof", {{1, 1}, {2, 2}, {2, 2}}},
    {"curv", {{1, 1}, {1, 1}, {1, 1}}},   {"dvof", {{0, 1}, {1, 2}, {1, 2}}},
    {"flxx", {{0, 1}, {0, 0}, {0, 0}}},   {"flxy", {{0, 0}, {0, 1}, {0, 0}}},
    {"flxz", {{0, 0}, {0, 0}, {0, 1}}},   {"normal", {{0, 0}, {1, 1}, {1, 1}}},
    {"src", {{0, 0}, {0, 0}, {0, 0}}},
  };
  Nadds r{};
  for (const E & e : tab)
    if (!strcmp(e.n, name)) {
      memcpy(r.v, e.a, sizeof r.v);
      return r;
    }
  return r;
}

static int array_prepare(const domain_t * domain, const char * name, size_t elsize, array_t * a) {
  vof_ctx * ctx = ctx_of(domain);
  const Nadds na = nadds_of(name);
  size_t nitems = 1;
  a->nadds = (int (*)[2])calloc(NDIMS, sizeof(int[2]));
  if (!a->nadds) return fail("calloc failed");
  for (int d = 0; d < NDIMS; d++) {
    nitems *= domain->mysizes[d] + na.v[d][0] + na.v[d][1];
    a->nadds[d][0] = na.v[d][0];
    a->nadds[d][1] = na.v[d][1];
  }
  a->size = elsize;
  a->datasize = nitems * elsize;
  a->data = nullptr;
  HIPCHK(hipMalloc(&a->data, a->datasize));
  HIPCHK(hipMemsetAsync(a->data, 0, a->datasize, ctx->stream));
  return 0;
}

static int array_free(array_t * a) {
  if (a->data) dev_free(a->data);
  free(a->nadds);
  a->data = nullptr;
  a->nadds = nullptr;
  return 0;
}

static size_t padded_x(const array_t * a, const domain_t * d) {
  return d->mysizes[0] + a->nadds[0][0] + a->nadds[0][1];
}

// array.load / array.dump (src/array.c:80-206): the NPY holds the interior
// y (z) rows with the x padding
static int array_load(const domain_t * domain, const char * dir, const char * name, array_t * a) {
  vof_ctx * ctx = ctx_of(domain);
  const size_t n0 = padded_x(a, domain);
  std::vector<size_t> shape;
#if NDIMS == 3
  shape.push_back(domain->glsizes[2]);
#endif
  shape.push_back(domain->glsizes[1]);
  shape.push_back(n0);
  size_t nrow = domain->glsizes[1];
#if NDIMS == 3
  nrow *= domain->glsizes[2];
#endif
  std::vector<double> buf(n0 * nrow);
  if (npy::read(dir, name, kNpyDouble, shape, buf.data(), sizeof(double))) return 1;
  std::vector<double> full(a->datasize / sizeof(double), 0.);
  HIPCHK(hipMemcpyAsync(full.data(), a->data, a->datasize, hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  const size_t n1 = domain->mysizes[1] + a->nadds[1][0] + a->nadds[1][1];
  const int hy = a->nadds[1][0];
#if NDIMS == 3
  const int hz = a->nadds[2][0];
  const size_t nzz = domain->glsizes[2];
#else
  const int hz = 0;
  const size_t nzz = 1;
#endif
  for (size_t k = 0; k < nzz; k++)
    for (size_t j = 0; j < domain->glsizes[1]; j++)
      memcpy(&full[n0 * ((j + hy) + n1 * (k + hz))], &buf[n0 * (j + domain->glsizes[1] * k)],
             n0 * sizeof(double));
  HIPCHK(hipMemcpyAsync(a->data, full.data(), a->datasize, hipMemcpyHostToDevice, ctx->stream));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  return 0;
}

static int array_dump(const domain_t * domain, const char * dir, const char * name, const array_t * a) {
  vof_ctx * ctx = ctx_of(domain);
  const size_t n0 = padded_x(a, domain);
  std::vector<double> full(a->datasize / sizeof(double));
  HIPCHK(hipMemcpyAsync(full.data(), a->data, a->datasize, hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  std::vector<size_t> shape;
#if NDIMS == 3
  shape.push_back(domain->glsizes[2]);
#endif
  shape.push_back(domain->glsizes[1]);
  shape.push_back(n0);
  size_t nrow = domain->glsizes[1];
#if NDIMS == 3
  nrow *= domain->glsizes[2];
  const int hz = a->nadds[2][0];
  const size_t nzz = domain->glsizes[2];
#else
  const int hz = 0;
  const size_t nzz = 1;
#endif
  std::vector<double> buf(n0 * nrow);
  const size_t n1 = domain->mysizes[1] + a->nadds[1][0] + a->nadds[1][1];
  const int hy = a->nadds[1][0];
  for (size_t k = 0; k < nzz; k++)
    for (size_t j = 0; j < domain->glsizes[1]; j++)
      memcpy(&buf[n0 * (j + domain->glsizes[1] * k)], &full[n0 * ((j + hy) + n1 * (k + hz))],
             n0 * sizeof(double));
  return npy::write(dir, name, kNpyDouble, shape, buf.data(), sizeof(double));
}

extern "C" int vof_array_to_host(const domain_t * domain, const array_t * a, void * host) {
  CTX_OR_FAIL(domain);
  if (!a || !a->data || !host) return fail("vof_array_to_host: null argument");
  HIPCHK(hipMemcpyAsync(host, a->data, a->datasize, hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  return 0;
}

extern "C" int vof_array_from_host(const domain_t * domain, array_t * a, const void * host) {
  CTX_OR_FAIL(domain);
  if (!a || !a->data || !host) return fail("vof_array_from_host: null argument");
  HIPCHK(hipMemcpyAsync(a->data, host, a->datasize, hipMemcpyHostToDevice, ctx->stream));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  return 0;
}

extern "C" int vof_synchronize(const domain_t * domain) {
  CTX_OR_FAIL(domain);
  HIPCHK(hipStreamSynchronize(ctx->stream));
  return 0;
}

// device memory for callers of the tdm vtable and linear_system that hold no
// HIP runtime of their own: the buffers come from this library's runtime
extern "C" int vof_device_alloc(void ** ptr, size_t bytes) {
  if (!ptr) return fail("vof_device_alloc: ptr is NULL");
  *ptr = nullptr;
  HIPCHK(hipMalloc(ptr, bytes ? bytes : 1));
  return 0;
}

extern "C" int vof_device_free(void * ptr) {
  if (ptr) HIPCHK(hipFree(ptr));
  return 0;
}

extern "C" int vof_device_copy(void * dst, const void * src, size_t bytes, int kind) {
  if (!dst || !src) return fail("vof_device_copy: null pointer");
  hipMemcpyKind k;
  switch (kind) {
    case 1: k = hipMemcpyHostToDevice; break;
    case 2: k = hipMemcpyDeviceToHost; break;
    case 3: k = hipMemcpyDeviceToDevice; break;
    default: return fail("vof_device_copy: kind %d is not 1 (to device), 2 (to host) or 3 (device to device)", kind);
  }
  HIPCHK(hipMemcpy(dst, src, bytes, k));
  return 0;
}

// ---------------------------------------------------------------------------
// FFT plans and twiddles

static int make_plan(int n, FftPlan * P) {
  P->n = n;
  P->nst = 0;
  int m = n;
  const int rads[4] = {4, 2, 3, 5};
  for (int r : rads) {
    while (m % r == 0) {
      if (P->nst >= 16) return fail("FFT length %d has too many factors", n);
      P->rad[P->nst++] = r;
      m /= r;
    }
  }
  if (m != 1) return fail("FFT length %d must factor into 2, 3 and 5", n);
  if (n > vofk::kFftElems) return fail("FFT length %d exceeds %d", n, vofk::kFftElems);
  return 0;
}

// (cos, sin)(2 pi m / n) reduced to the first octant so that symmetric
// angles give identical magnitudes and the axes are exact
static std::vector<double2> make_twiddles(int n) {
  std::vector<double2> tw(n);
  const double pi = 3.14159265358979324;
  for (int k = 0; k < n; k++) {
    const long m8 = 8L * k;
    double c, s;
    if (m8 % n == 0) {
      static const double cs[8][2] = {
        {1., 0.}, {0.70710678118654752440, 0.70710678118654752440}, {0., 1.},
        {-0.70710678118654752440, 0.70710678118654752440}, {-1., 0.},
        {-0.70710678118654752440, -0.70710678118654752440}, {0., -1.},
        {0.70710678118654752440, -0.70710678118654752440},
      };
      const int oct = (int)(m8 / n);
      c = cs[oct][0];
      s = cs[oct][1];
    } else {
      const long kk = 4L * k;
      const int q = (int)(kk / n);
      const long rem = kk - (long)q * n;
      double cr, sr;
      if (2 * rem <= n) {
        const double r = 0.5 * pi * (double)rem / n;
        cr = cos(r); sr = sin(r);
      } else {
        const double r = 0.5 * pi * (double)(n - rem) / n;
        cr = sin(r); sr = cos(r);
      }
      switch (q) {
        case 0: c = cr; s = sr; break;
        case 1: c = -sr; s = cr; break;
        case 2: c = -cr; s = -sr; break;
        default: c = sr; s = -cr; break;
      }
    }
    tw[k] = make_double2(c, s);
  }
  return tw;
}

// ---------------------------------------------------------------------------
// domain

static int ctx_create(domain_t * domain) {
  vof_ctx * ctx = new vof_ctx();
  domain->info = ctx;
  ctx->nx = (int)domain->glsizes[0];
  ctx->ny = (int)domain->glsizes[1];
#if NDIMS == 3
  ctx->nz = (int)domain->glsizes[2];
  ctx->dz = domain->dz;
#else
  ctx->nz = 1;
  ctx->dz = 0.;
#endif
  ctx->dy = domain->dy;
  ctx->schedule = 1;
  // the one-kernel predictor (F_PRED) is opt-in: with two work-groups per
  // CU (64 KiB of LDS each) its serial line sweeps and field update are not
  // hidden behind the source stencils (256^3: 1.47 ms against 0.70 + 0.50
  // ms for the two-kernel predictor)
  ctx->fuse = 63;
  ctx->fft_elems = 1024;
  ctx->fft_radix = 16;
  ctx->thomas_rows = 8;
  ctx->thomas_part = 1;
  // two streams pay from about 2^20 cells; below, the launches of a step
  // are short and the cross-stream events cost more than the overlap gains
  ctx->iface_overlap = (size_t)ctx->nx * (size_t)ctx->ny * (size_t)ctx->nz >= ((size_t)1 << 20) ? 1 : 0;
  ctx->poisson_refine = -1;
  ctx->adj_floor = 1e-2;
  ctx->lines_lpb = 0;
  ctx->rhs_rows = 8;
  ctx->iface_planes = 8;
  ctx->d_vof_spare = nullptr;
  ctx->vof_bytes = 0;
  for (int m = 0; m < 4; m++) ctx->d_fspare[m] = nullptr;
  ctx->pred_rows = 8;
  ctx->freeze_mode = 0;
  ctx->freeze_pos = ctx->freeze_n = 0;
  ctx->prof_on = false;
  ctx->ev_next = 0;
  const int nx = ctx->nx, ny = ctx->ny, nz = ctx->nz;
  if (nx < 4 || ny < 4 || (NDIMS == 3 && nz < 4))
    return fail("grid %dx%dx%d too small (at least 4 cells per direction)", nx, ny, nz);
  // reference compile-time parameters (src/param/*.c)
  for (int m = 0; m < P_COUNT; m++) { ctx->prm[m] = 0.; ctx->prm_set[m] = false; }
  ctx->prm[P_BUOY] = 1.;  ctx->prm_set[P_BUOY] = true;
  ctx->prm[P_UY_XM] = 0.; ctx->prm_set[P_UY_XM] = true;
  ctx->prm[P_UY_XP] = 0.; ctx->prm_set[P_UY_XP] = true;
  ctx->prm[P_UZ_XM] = 0.; ctx->prm_set[P_UZ_XM] = true;
  ctx->prm[P_UZ_XP] = 0.; ctx->prm_set[P_UZ_XP] = true;
  ctx->prm[P_T_XM] = +0.5; ctx->prm_set[P_T_XM] = true;
  ctx->prm[P_T_XP] = -0.5; ctx->prm_set[P_T_XP] = true;
  HIPCHK(hipStreamCreate(&ctx->stream));
  // metrics
  HIPCHK(hipMalloc(&ctx->d_dxf, sizeof(double) * nx));
  HIPCHK(hipMalloc(&ctx->d_dxc, sizeof(double) * (nx + 1)));
  HIPCHK(hipMemcpy(ctx->d_dxf, domain->dxf, sizeof(double) * nx, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(ctx->d_dxc, domain->dxc, sizeof(double) * (nx + 1), hipMemcpyHostToDevice));
  ctx->grid = Grid{nx, ny, nz, ctx->dy, ctx->dz, ctx->d_dxf, ctx->d_dxc, 0,
                   0., 0., 0., 0., 0., 0., 0., 0., nullptr};
  {
    // the stencils' quotients, by the divisions the reference spells
    Grid & g = ctx->grid;
    const double dy = ctx->dy, dz = ctx->dz;
    g.hdy = +0.5 / dy;
    g.rdy = 1. / dy;
    g.lyy = +1. / dy / dy;
    g.cyy = -2. / dy / dy;
    g.hdz = +0.5 / dz;
    g.rdz = 1. / dz;
    g.lzz = +1. / dz / dz;
    g.czz = -2. / dz / dz;
    std::vector<double> xt((size_t)vofk::kXT * (nx + 2), 0.);
    const double * dxf = domain->dxf;
    const double * dxc = domain->dxc;
    for (int i = 1; i <= nx + 1; i++) {
      double * e = &xt[(size_t)vofk::kXT * i];
      e[vofk::XT_HXC] = +0.5 / dxc[i - 1];
      e[vofk::XT_RXC] = 1. / dxc[i - 1];
      if (i >= 2) e[vofk::XT_WXM] = 0.5 * dxf[i - 2] / dxc[i - 1];
      if (i <= nx) {
        e[vofk::XT_WXP] = 0.5 * dxf[i - 1] / dxc[i - 1];
        e[vofk::XT_HXF] = +0.5 / dxf[i - 1];
        e[vofk::XT_RXF] = 1. / dxf[i - 1];
      }
    }
    double * d = nullptr;
    HIPCHK(hipMalloc(&d, sizeof(double) * xt.size()));
    HIPCHK(hipMemcpy(d, xt.data(), sizeof(double) * xt.size(), hipMemcpyHostToDevice));
    g.xt = d;
  }
  // Laplacians in x (predict/ux.c:35-62, uy.c:35-62, t.c:33-60)
  const double * dxf = domain->dxf;
  const double * dxc = domain->dxc;
#define DXF(I) (dxf[(I) - 1])
#define DXC(I) (dxc[(I) - 1])
  ctx->lapx_ux.resize(3 * (nx - 1));
  for (int i = 2; i <= nx; i++) {
    const double l = 1. / DXF(i - 1) / DXC(i);
    const double u = 1. / DXF(i) / DXC(i);
    const double c = -l - u;
    ctx->lapx_ux[3 * (i - 2) + 0] = l;
    ctx->lapx_ux[3 * (i - 2) + 1] = c;
    ctx->lapx_ux[3 * (i - 2) + 2] = u;
  }
  ctx->lapx_c.resize(3 * nx);
  for (int i = 1; i <= nx; i++) {
    const double l = 1. / DXC(i) / DXF(i);
    const double u = 1. / DXC(i + 1) / DXF(i);
    const double c = -l - u;
    ctx->lapx_c[3 * (i - 1) + 0] = l;
    ctx->lapx_c[3 * (i - 1) + 1] = c;
    ctx->lapx_c[3 * (i - 1) + 2] = u;
  }
  HIPCHK(hipMalloc(&ctx->d_lapx_ux, sizeof(double) * 3 * (nx - 1)));
  HIPCHK(hipMalloc(&ctx->d_lapx_c, sizeof(double) * 3 * nx));
  HIPCHK(hipMemcpy(ctx->d_lapx_ux, ctx->lapx_ux.data(), sizeof(double) * 3 * (nx - 1), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(ctx->d_lapx_c, ctx->lapx_c.data(), sizeof(double) * 3 * nx, hipMemcpyHostToDevice));
  // Poisson operator in x (dft.c:160-190) and eigenvalues (dft.c:263-291)
  {
    std::vector<double> l(nx), u(nx);
    for (int i = 1; i <= nx; i++) {
      l[i - 1] = 1. / DXC(i) / DXF(i);
      u[i - 1] = 1. / DXC(i + 1) / DXF(i);
    }
    HIPCHK(hipMalloc(&ctx->d_pl, sizeof(double) * nx));
    HIPCHK(hipMalloc(&ctx->d_pu, sizeof(double) * nx));
    HIPCHK(hipMemcpy(ctx->d_pl, l.data(), sizeof(double) * nx, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(ctx->d_pu, u.data(), sizeof(double) * nx, hipMemcpyHostToDevice));
    // elimination factors of the (ky, kz) = (0, 0) line, whose eigenvalue is
    // 0 in every step: the operations of k_thomas_spec's sweep for it, once
    // (the partitioned solve eliminates that line with them)
    std::vector<double> st(3 + 2 * (size_t)nx, 0.);
    if (nx >= 2) {
      const double eval = 0.;
      double cc = -l[0] - u[0] + eval;
      cc += l[0];
      st[0] = cc;
      double vprev = u[0] / cc;
      st[3 + nx] = vprev;
      for (int i = 1; i < nx - 1; i++) {
        const double c = -l[i] - u[i] + eval;
        const double val = 1. / (c - l[i] * vprev);
        const double vi = val * (u[i]);
        st[3 + i] = val;
        st[3 + nx + i] = vi;
        vprev = vi;
      }
      const int i = nx - 1;
      double c = -l[i] - u[i] + eval;
      c += u[i];
      const double val = c - l[i] * vprev;
      st[2] = fabs(val) > DBL_EPSILON ? 1. : 0.;
      st[1] = st[2] != 0. ? 1. / val : 0.;
    }
    HIPCHK(hipMalloc(&ctx->d_stab, sizeof(double) * st.size()));
    HIPCHK(hipMemcpy(ctx->d_stab, st.data(), sizeof(double) * st.size(), hipMemcpyHostToDevice));
  }
#undef DXF
#undef DXC
  const double g_pi = 3.14159265358979324;
  ctx->nh = ny / 2 + 1;
  {
    std::vector<double> ey(ctx->nh);
    const double gy = domain->lengths[1] / ny;
    for (int m = 0; m < ctx->nh; m++) {
      const double s = sin(g_pi * m / (double)ny);
      ey[m] = -4. / (gy * gy) * (s * s);
    }
    HIPCHK(hipMalloc(&ctx->d_ey, sizeof(double) * ctx->nh));
    HIPCHK(hipMemcpy(ctx->d_ey, ey.data(), sizeof(double) * ctx->nh, hipMemcpyHostToDevice));
  }
  {
    std::vector<double> ez(nz, 0.);
#if NDIMS == 3
    const double gz = domain->lengths[2] / nz;
    for (int m = 0; m < nz; m++) {
      const double s = sin(g_pi * m / (double)nz);
      ez[m] = -4. / (gz * gz) * (s * s);
    }
#endif
    HIPCHK(hipMalloc(&ctx->d_ez, sizeof(double) * nz));
    HIPCHK(hipMemcpy(ctx->d_ez, ez.data(), sizeof(double) * nz, hipMemcpyHostToDevice));
  }
  if (make_plan(ny, &ctx->plan_y)) return 1;
  {
    std::vector<double2> tw = make_twiddles(ny);
    HIPCHK(hipMalloc(&ctx->d_twy, sizeof(double2) * ny));
    HIPCHK(hipMemcpy(ctx->d_twy, tw.data(), sizeof(double2) * ny, hipMemcpyHostToDevice));
  }
#if NDIMS == 3
  if (make_plan(nz, &ctx->plan_z)) return 1;
  {
    std::vector<double2> tw = make_twiddles(nz);
    HIPCHK(hipMalloc(&ctx->d_twz, sizeof(double2) * nz));
    HIPCHK(hipMemcpy(ctx->d_twz, tw.data(), sizeof(double2) * nz, hipMemcpyHostToDevice));
  }
#else
  ctx->d_twz = nullptr;
#endif
  const size_t ncell = (size_t)nx * ny * nz;
  const size_t nline = (size_t)ctx->nh * nz;
  HIPCHK(hipMalloc(&ctx->d_spec, sizeof(double2) * nx * nline));
  HIPCHK(hipMalloc(&ctx->d_vbuf, sizeof(double) * nx * nline));
  ctx->vbuf_ready = false;
  HIPCHK(hipMalloc(&ctx->d_delta, sizeof(double) * ncell * (NDIMS + 1)));
  ctx->npart = 8192;
  HIPCHK(hipMalloc(&ctx->d_part, sizeof(double) * ctx->npart * vofk::kDiagN));
  HIPCHK(hipHostMalloc(&ctx->h_scalar, sizeof(double) * 64));
  // Thomas tables: header 4 + l, val, v
  ctx->tab_stride = 4 + 3 * (size_t)nx;
  ctx->tab_rk = 4 * ctx->tab_stride;
  HIPCHK(hipMalloc(&ctx->d_tab, sizeof(double) * ctx->tab_stride * 12));
  HIPCHK(hipHostMalloc(&ctx->h_tab, sizeof(double) * ctx->tab_stride * 12));
  for (int m = 0; m < 12; m++) HIPCHK(hipEventCreateWithFlags(&ctx->tab_done[m], hipEventDisableTiming));
  return 0;
}

extern "C" int domain_init(const char dirname_ic[], domain_t * domain) {
  if (!domain) return fail("domain_init: domain is NULL");
  size_t gl[NDIMS];
  double len[NDIMS];
  if (npy::read(dirname_ic, "glsizes", kNpySizeT, {(size_t)NDIMS}, gl, sizeof(size_t))) return 1;
  if (npy::read(dirname_ic, "lengths", kNpyDouble, {(size_t)NDIMS}, len, sizeof(double))) return 1;
  for (int d = 0; d < NDIMS; d++) {
    domain->glsizes[d] = gl[d];
    domain->mysizes[d] = gl[d];
    domain->offsets[d] = 0;
    domain->lengths[d] = len[d];
  }
  const size_t nx = gl[0];
  domain->xf = (double *)calloc(nx + 1, sizeof(double));
  domain->xc = (double *)calloc(nx + 2, sizeof(double));
  domain->dxf = (double *)calloc(nx, sizeof(double));
  domain->dxc = (double *)calloc(nx + 1, sizeof(double));
  if (npy::read(dirname_ic, "xf", kNpyDouble, {nx + 1}, domain->xf, sizeof(double))) return 1;
  if (npy::read(dirname_ic, "xc", kNpyDouble, {nx + 2}, domain->xc, sizeof(double))) return 1;
  // domain.c:103-143
  for (size_t i = 1; i <= nx; i++) domain->dxf[i - 1] = domain->xf[i] - domain->xf[i - 1];
  for (size_t i = 1; i <= nx + 1; i++) domain->dxc[i - 1] = domain->xc[i] - domain->xc[i - 1];
  domain->dy = domain->lengths[1] / domain->glsizes[1];
#if NDIMS == 3
  domain->dz = domain->lengths[2] / domain->glsizes[2];
#endif
  if (ctx_create(domain)) return 1;
  bool uni = false;
  domain_check_x_grid_is_uniform(domain, &uni);
  ctx_of(domain)->x_uniform = uni;
  return 0;
}

extern "C" int domain_save(const char dirname[], const domain_t * domain) {
  if (!domain) return fail("domain_save: domain is NULL");
  const size_t nx = domain->glsizes[0];
  if (npy::write(dirname, "glsizes", kNpySizeT, {(size_t)NDIMS}, domain->glsizes, sizeof(size_t))) return 1;
  if (npy::write(dirname, "lengths", kNpyDouble, {(size_t)NDIMS}, domain->lengths, sizeof(double))) return 1;
  if (npy::write(dirname, "xf", kNpyDouble, {nx + 1}, domain->xf, sizeof(double))) return 1;
  if (npy::write(dirname, "xc", kNpyDouble, {nx + 2}, domain->xc, sizeof(double))) return 1;
  return 0;
}

extern "C" int domain_check_x_grid_is_uniform(const domain_t * domain, bool * x_grid_is_uniform) {
  // domain.c:75-98
  const size_t nx = domain->glsizes[0];
  double mn = +1. * DBL_MAX, mx = -1. * DBL_MAX;
  for (size_t i = 0; i < nx; i++) {
    mn = fmin(mn, domain->dxf[i]);
    mx = fmax(mx, domain->dxf[i]);
  }
  *x_grid_is_uniform = fabs(mx - mn) < 1.e-15;
  return 0;
}

extern "C" int domain_finalise(domain_t * domain) {
  vof_ctx * ctx = ctx_of(domain);
  if (ctx) {
    (void)hipStreamSynchronize(ctx->stream);
    dev_free((void *)ctx->grid.xt);
    dev_free(ctx->d_dxf); dev_free(ctx->d_dxc); dev_free(ctx->d_lapx_ux); dev_free(ctx->d_lapx_c);
    // tables inside a live batch's block belong to the batch
    if (!ctx->tab_shared) {
      dev_free(ctx->d_tab);
      host_free(ctx->h_tab);
    }
    for (int m = 0; m < 12; m++) (void)hipEventDestroy(ctx->tab_done[m]);
    dev_free(ctx->d_twy); if (ctx->d_twz) dev_free(ctx->d_twz);
    dev_free(ctx->d_pl); dev_free(ctx->d_pu); dev_free(ctx->d_ey); dev_free(ctx->d_ez); dev_free(ctx->d_stab);
    dev_free(ctx->d_spec32); dev_free(ctx->d_twy32); dev_free(ctx->d_twz32);
    dev_free(ctx->d_spec); dev_free(ctx->d_vbuf); dev_free(ctx->d_delta); dev_free(ctx->d_part);
    if (ctx->d_vof_spare) dev_free(ctx->d_vof_spare);
    for (int m = 0; m < 4; m++) dev_free(ctx->d_fspare[m]);
    for (void * b : ctx->freeze_slots) dev_free(b);
    for (int m = 0; m < 9; m++) dev_free(ctx->d_stat[m]);
    host_free(ctx->h_scalar);
    prof_flush(ctx);
    for (auto & e : ctx->ev_pool) (void)hipEventDestroy(e);
    if (ctx->side) {
      (void)hipStreamSynchronize(ctx->side);
      (void)hipEventDestroy(ctx->ev_rhs);
      (void)hipEventDestroy(ctx->ev_iface);
      (void)hipEventDestroy(ctx->ev_main);
      (void)hipStreamDestroy(ctx->side);
    }
    (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    domain->info = nullptr;
  }
  free(domain->xf); free(domain->xc); free(domain->dxf); free(domain->dxc);
  domain->xf = domain->xc = domain->dxf = domain->dxc = nullptr;
  return 0;
}

extern "C" int vof_set_param(domain_t * domain, const char name[], double value) {
  CTX_OR_FAIL(domain);
  for (int m = 0; m < P_COUNT; m++)
    if (!strcmp(name, kParamNames[m])) {
      ctx->prm[m] = value;
      ctx->prm_set[m] = true;
      return 0;
    }
  return fail("vof_set_param: unknown parameter %s", name);
}

extern "C" int vof_get_param(const domain_t * domain, const char name[], double * value) {
  CTX_OR_FAIL(domain);
  for (int m = 0; m < P_COUNT; m++)
    if (!strcmp(name, kParamNames[m])) return param_get(ctx, m, value);
  return fail("vof_get_param: unknown parameter %s", name);
}

extern "C" int vof_set_schedule(domain_t * domain, int schedule) {
  CTX_OR_FAIL(domain);
  if (schedule < 0 || schedule > 1) return fail("vof_set_schedule: %d is not 0 or 1", schedule);
  ctx->schedule = schedule;
  return 0;
}

// ---------------------------------------------------------------------------
// fluid

#if NDIMS == 2
#define FOR_FLUID_ARRAYS(X) X(ux) X(uy) X(p) X(psi) X(t)
#else
#define FOR_FLUID_ARRAYS(X) X(ux) X(uy) X(uz) X(p) X(psi) X(t)
#endif

extern "C" int fluid_init(const char dirname_ic[], const domain_t * domain, fluid_t * fluid) {
  CTX_OR_FAIL(domain);
#define PREP(n) if (array_prepare(domain, #n, sizeof(double), &fluid->n)) return 1;
  FOR_FLUID_ARRAYS(PREP)
#undef PREP
  for (int m = 0; m < 3; m++) {
    if (array_prepare(domain, "srcux", sizeof(double), &fluid->srcux[m])) return 1;
    if (array_prepare(domain, "srcuy", sizeof(double), &fluid->srcuy[m])) return 1;
#if NDIMS == 3
    if (array_prepare(domain, "srcuz", sizeof(double), &fluid->srcuz[m])) return 1;
#endif
    if (array_prepare(domain, "srct", sizeof(double), &fluid->srct[m])) return 1;
  }
  if (array_load(domain, dirname_ic, "ux", &fluid->ux)) return 1;
  if (array_load(domain, dirname_ic, "uy", &fluid->uy)) return 1;
#if NDIMS == 3
  if (array_load(domain, dirname_ic, "uz", &fluid->uz)) return 1;
#endif
  if (array_load(domain, dirname_ic, "p", &fluid->p)) return 1;
  if (array_load(domain, dirname_ic, "t", &fluid->t)) return 1;
  if (fluid_update_boundaries_ux(domain, &fluid->ux)) return 1;
  if (fluid_update_boundaries_uy(domain, &fluid->uy)) return 1;
#if NDIMS == 3
  if (fluid_update_boundaries_uz(domain, &fluid->uz)) return 1;
#endif
  if (fluid_update_boundaries_p(domain, &fluid->p)) return 1;
  if (fluid_update_boundaries_t(domain, &fluid->t)) return 1;
  // fluid/init.c:90-93
  if (param_get(ctx, P_PR, &fluid->Pr)) return 1;
  if (param_get(ctx, P_RA, &fluid->Ra)) return 1;
  fluid->m_dif = 1. * sqrt(fluid->Pr) / sqrt(fluid->Ra);
  fluid->t_dif = 1. / sqrt(fluid->Pr) / sqrt(fluid->Ra);
  HIPCHK(hipStreamSynchronize(ctx->stream));
  return 0;
}

extern "C" int fluid_finalise(const domain_t * domain, fluid_t * fluid) {
  if (ctx_of(domain)) (void)hipStreamSynchronize(ctx_of(domain)->stream);
#define FREE(n) array_free(&fluid->n);
  FOR_FLUID_ARRAYS(FREE)
#undef FREE
  for (int m = 0; m < 3; m++) {
    array_free(&fluid->srcux[m]);
    array_free(&fluid->srcuy[m]);
#if NDIMS == 3
    array_free(&fluid->srcuz[m]);
#endif
    array_free(&fluid->srct[m]);
  }
  return 0;
}

extern "C" int fluid_save(const char dirname[], const domain_t * domain, const fluid_t * fluid) {
  CTX_OR_FAIL(domain);
  (void)ctx;
  if (npy::write(dirname, "m_dif", kNpyDouble, {}, &fluid->m_dif, sizeof(double))) return 1;
  if (npy::write(dirname, "t_dif", kNpyDouble, {}, &fluid->t_dif, sizeof(double))) return 1;
  if (array_dump(domain, dirname, "ux", &fluid->ux)) return 1;
  if (array_dump(domain, dirname, "uy", &fluid->uy)) return 1;
#if NDIMS == 3
  if (array_dump(domain, dirname, "uz", &fluid->uz)) return 1;
#endif
  if (array_dump(domain, dirname, "p", &fluid->p)) return 1;
  if (array_dump(domain, dirname, "t", &fluid->t)) return 1;
  return 0;
}

extern "C" int fluid_update_boundaries_ux(const domain_t * domain, array_t * ux) {
  CTX_OR_FAIL(domain);
  return bc_ux(ctx, (double *)ux->data);
}
extern "C" int fluid_update_boundaries_uy(const domain_t * domain, array_t * uy) {
  CTX_OR_FAIL(domain);
  return bc_uy(ctx, (double *)uy->data);
}
#if NDIMS == 3
extern "C" int fluid_update_boundaries_uz(const domain_t * domain, array_t * uz) {
  CTX_OR_FAIL(domain);
  return bc_uz(ctx, (double *)uz->data);
}
#endif
extern "C" int fluid_update_boundaries_p(const domain_t * domain, array_t * p) {
  CTX_OR_FAIL(domain);
  return bc_neumann(ctx, (double *)p->data);
}
extern "C" int fluid_update_boundaries_psi(const domain_t * domain, array_t * psi) {
  CTX_OR_FAIL(domain);
  return bc_neumann(ctx, (double *)psi->data);
}
extern "C" int fluid_update_boundaries_t(const domain_t * domain, array_t * t) {
  CTX_OR_FAIL(domain);
  return bc_t(ctx, (double *)t->data);
}

// 1 for the first spectral Thomas solve (which writes the back-substitution
// factors), 0 afterwards
static int vbuf_store(vof_ctx * ctx) {
  const int first = ctx->vbuf_ready ? 0 : 1;
  ctx->vbuf_ready = true;
  return first;
}

// fluid_decide_dt (decide_dt.c:104-145)
static int decide_dt_enqueue(vof_ctx * ctx, const fluid_t * fluid) {
  // 64 x 4 (i, j) tiles, z planes split over gz work-groups so that the
  // partial minima fit d_part
  const unsigned gx = (unsigned)((ctx->nx + 63) / 64);
  unsigned gy = (unsigned)((ctx->ny + 3) / 4);
  unsigned gz = NDIMS == 3 ? (unsigned)((ctx->nz + 15) / 16) : 1u;
  while (gz > 1 && (size_t)gx * gy * gz > (size_t)ctx->npart) gz--;
  while (gy > 1 && (size_t)gx * gy * gz > (size_t)ctx->npart) gy--;
  if ((size_t)gx * gy * gz > (size_t)ctx->npart) return fail("decide_dt: %u x-tiles exceed the partial buffer", gx);
  const int nb = (int)(gx * gy * gz);
#if NDIMS == 3
  const double * uz = (const double *)fluid->uz.data;
#else
  const double * uz = nullptr;
#endif
  KLB(ctx, "k_dt_partial", vofk::kRedThreads, vofk::k_dt_partial_body, vofk::k_dt_partial, dim3(gx, gy, gz),
      dim3(vofk::kRedThreads), 0, ctx->grid, (const double *)fluid->ux.data, (const double *)fluid->uy.data,
      uz, ctx->d_part);
  KLB(ctx, "k_min_final", vofk::kRedThreads, vofk::k_min_final_body, vofk::k_min_final, dim3(1),
      dim3(vofk::kRedThreads), 0, (const double *)ctx->d_part, nb, ctx->d_part + ctx->npart);
  // a batch gathers the minima of all its cases itself
  if (!ctx->rec)
    HIPCHK(hipMemcpyAsync(ctx->h_scalar, ctx->d_part + ctx->npart, sizeof(double),
                          hipMemcpyDeviceToHost, ctx->stream));
  return 0;
}

static int decide_dt_finish(vof_ctx * ctx, const domain_t * domain, const fluid_t * fluid, double * dt,
                            bool sync = true) {
  if (sync) HIPCHK(hipStreamSynchronize(ctx->stream));
  double coef_adv = 0., coef_dif = 0.;
  if (param_get(ctx, P_COEF_ADV, &coef_adv)) return 1;
  if (param_get(ctx, P_COEF_DIF, &coef_dif)) return 1;
  double res = ctx->h_scalar[0];
  res *= coef_adv;
  // decide_dt_dif: only y (and z) are explicit (param/implicit.c:3-7)
  double hs[3];
  hs[0] = DBL_MAX;
  for (int i = 2; i <= ctx->nx; i++) hs[0] = fmin(hs[0], domain->dxc[i - 1]);
  hs[1] = ctx->dy;
  hs[2] = ctx->dz;
  const double difs[2] = {fluid->m_dif, fluid->t_dif};
  for (int f = 0; f < 2; f++)
    for (int d = 1; d < NDIMS; d++) {
      const double v = coef_dif / difs[f] * 0.5 / NDIMS * (hs[d] * hs[d]);
      res = fmin(res, v);
    }
  *dt = res;
  return 0;
}

extern "C" int fluid_decide_dt(const domain_t * domain, const fluid_t * fluid, double * dt) {
  CTX_OR_FAIL(domain);
  if (decide_dt_enqueue(ctx, fluid)) return 1;
  return decide_dt_finish(ctx, domain, fluid, dt);
}

// runge_kutta.c:8-15
static const double kRk[3][3] = {
  {+32. / 60., 0. / 60., +32. / 60. + 0. / 60.},
  {+25. / 60., -17. / 60., +25. / 60. + -17. / 60.},
  {+45. / 60., -25. / 60., +45. / 60. + -25. / 60.},
};

// Thomas factors shared by every line of one predictor solve (tdm.c:17-50
// with the coefficients of solve_in_x, ux.c:251-272)
static void build_tab(double * tab, int n, double prefactor, const double * lapx) {
  double * L = tab + 4;
  double * VAL = tab + 4 + n;
  double * V = tab + 4 + 2 * n;
  std::vector<double> l(n), c(n), u(n);
  for (int i = 0; i < n; i++) {
    l[i] = -prefactor * lapx[3 * i + 0];
    c[i] = 1. - prefactor * lapx[3 * i + 1];
    u[i] = -prefactor * lapx[3 * i + 2];
  }
  for (int i = 0; i < n; i++) L[i] = l[i];
  V[0] = u[0] / c[0];
  VAL[0] = 0.;
  tab[0] = c[0];
  for (int i = 1; i < n - 1; i++) {
    const double val = 1. / (c[i] - l[i] * V[i - 1]);
    VAL[i] = val;
    V[i] = val * (u[i]);
  }
  const double val = c[n - 1] - l[n - 1] * V[n - 2];
  if (fabs(val) > DBL_EPSILON) {
    tab[1] = 1. / val;
    tab[2] = 1.;
  } else {
    tab[1] = 0.;
    tab[2] = 0.;
  }
  tab[3] = 0.;
  VAL[n - 1] = 0.;
  V[n - 1] = 0.;
}

// enqueue: tables for `field` (0 ux, 1 uy, 2 uz, 3 t) at `rkstep`, copied
// from a pinned slot once the previous copy from that slot has completed
static int upload_tab(vof_ctx * ctx, int field, int rkstep, double prefactor, const double ** dtab) {
  if (!ctx->h_tab || !ctx->d_tab)
    return fail("this case has no Thomas tables (their allocation failed when its batch was destroyed)");
  const int slot = field * 3 + rkstep;
  double * h = ctx->h_tab + ctx->tab_rk * rkstep + ctx->tab_stride * field;
  double * d = ctx->d_tab + ctx->tab_rk * rkstep + ctx->tab_stride * field;
  const int n = field == 0 ? ctx->nx - 1 : ctx->nx;
  if (ctx->rec) {
    // batch: the tables of all cases travel in one copy per stage
    build_tab(h, n, prefactor, field == 0 ? ctx->lapx_ux.data() : ctx->lapx_c.data());
    *dtab = d;
    return 0;
  }
  HIPCHK(hipEventSynchronize(ctx->tab_done[slot]));
  build_tab(h, n, prefactor, field == 0 ? ctx->lapx_ux.data() : ctx->lapx_c.data());
  HIPCHK(hipMemcpyAsync(d, h, sizeof(double) * ctx->tab_stride, hipMemcpyHostToDevice, ctx->stream));
  HIPCHK(hipEventRecord(ctx->tab_done[slot], ctx->stream));
  *dtab = d;
  return 0;
}

static void swap_data(array_t * a, array_t * b) {
  void * tmp = a->data;
  a->data = b->data;
  b->data = tmp;
}

// fluid_predict_field (predict/main.c:62-79): the sources of every field
// from the old fields (compute_rhs, :31-40), then increments, the implicit x
// solve and the update of every field (predict, :42-53; ux.c:306-388 etc.).
// fused = 1: the source kernels write the increment directly (srcg is not
// stored) and the line solver reads it.
// vof_old / curv non-null: the sources of all fields come from one fused
// kernel that forms the CSF force from them (fused schedule); images: the
// line solver writes the halo images itself instead of boundary launches.
static int launch_iface(vof_ctx * ctx, interface_t * interface);

// the next stage's interface kernel on the side stream, from this point of
// the main stream on (its readers of the normals and curvature are done)
static int side_iface(vof_ctx * ctx) {
  interface_t * itf = const_cast<interface_t *>(ctx->pending_iface);
  ctx->pending_iface = nullptr;
  HIPCHK(hipEventRecord(ctx->ev_rhs, ctx->stream));
  HIPCHK(hipStreamWaitEvent(ctx->side, ctx->ev_rhs, 0));
  hipStream_t main_stream = ctx->stream;
  ctx->stream = ctx->side;
  const int rc = launch_iface(ctx, itf);
  ctx->stream = main_stream;
  if (rc) return 1;
  HIPCHK(hipEventRecord(ctx->ev_iface, ctx->side));
  return 0;
}

// the second buffers of the one-kernel predictor, allocated on first use
// with the fields' contents (the kernel rewrites every entry of the buffer it
// writes, walls and halo images included)
static int ensure_field_spares(vof_ctx * ctx, fluid_t * fluid) {
  array_t * arr[4] = {&fluid->ux, &fluid->uy, nullptr, &fluid->t};
#if NDIMS == 3
  arr[2] = &fluid->uz;
#endif
  for (int m = 0; m < 4; m++) {
    if (!arr[m] || ctx->d_fspare[m]) continue;
    HIPCHK(hipMalloc(&ctx->d_fspare[m], arr[m]->datasize));
    HIPCHK(hipMemcpyAsync(ctx->d_fspare[m], arr[m]->data, arr[m]->datasize, hipMemcpyDeviceToDevice, ctx->stream));
  }
  return 0;
}

// rows per thread of the one-kernel predictor at this grid (0: the x lines
// of a work-group do not fit its 64 KiB of LDS; the two-kernel path runs)
static int pred_rows_for(const vof_ctx * ctx, int * xw) {
  const int nx = ctx->nx;
  int w = ((min(nx, 256) + 63) / 64) * 64;   // lanes along x: 64, 128 or 256
  if (w == 192) w = 256;
  const int groups = 256 / w;                // row groups of a work-group
  *xw = w;
  const size_t line = sizeof(double) * (size_t)nx * (NDIMS + 1);
  int rt = ctx->pred_rows;
  while (rt > 0 && line * (size_t)(groups * rt) > 65536) rt /= 2;
  return rt;
}

static int predict_impl(const domain_t * domain, size_t rkstep, double dt, fluid_t * fluid,
                        const interface_t * interface, int fused, const double * vof_old = nullptr,
                        int images = 0, bool one_kernel = false) {
  CTX_OR_FAIL(domain);
  if (rkstep > 2) return fail("rkstep %zu out of range", rkstep);
  // reset_srcs (predict/main.c:9-29): the a <-> b swap; a (and g) are fully
  // rewritten by the source kernels
  if (rkstep != 0) {
    swap_data(&fluid->srcux[0], &fluid->srcux[1]);
    swap_data(&fluid->srcuy[0], &fluid->srcuy[1]);
#if NDIMS == 3
    swap_data(&fluid->srcuz[0], &fluid->srcuz[1]);
#endif
    swap_data(&fluid->srct[0], &fluid->srct[1]);
  }
  const double ca = kRk[rkstep][0] * dt, cb = kRk[rkstep][1] * dt, cg = kRk[rkstep][2] * dt;
  const double pm = 0.5 * kRk[rkstep][2] * dt * fluid->m_dif;
  const double pt = 0.5 * kRk[rkstep][2] * dt * fluid->t_dif;
  const int nx = ctx->nx, ny = ctx->ny, nz = ctx->nz;
  const size_t ncell = (size_t)nx * ny * nz;
#if NDIMS == 3
  const double * uz = (const double *)fluid->uz.data;
  const double * ifrcz = (const double *)interface->ifrcz.data;
#else
  const double * uz = nullptr;
#endif
  vofk::RhsArgs base{};
  base.ux = (const double *)fluid->ux.data;
  base.uy = (const double *)fluid->uy.data;
  base.uz = uz;
  base.p = (const double *)fluid->p.data;
  base.t = (const double *)fluid->t.data;
  base.add_buoyancy = ctx->prm[P_BUOY] != 0.;
  base.ca = ca; base.cb = cb; base.cg = cg;
  struct FieldJob {
    const char * name;
    void (*kern)(Grid, vofk::RhsArgs);
    array_t * src;
    array_t * field;
    const double * ifrc;
    const double * lapx;
    double dif;
    int kind;    // 0 ux, 1 uy, 2 uz, 3 t
  };
  FieldJob jobs[4];
  int nj = 0;
  jobs[nj++] = {"k_rhs_ux", vofk::k_rhs_ux, fluid->srcux, &fluid->ux,
                (const double *)interface->ifrcx.data, ctx->d_lapx_ux, fluid->m_dif, 0};
  jobs[nj++] = {"k_rhs_uy", vofk::k_rhs_uy, fluid->srcuy, &fluid->uy,
                (const double *)interface->ifrcy.data, ctx->d_lapx_c, fluid->m_dif, 1};
#if NDIMS == 3
  jobs[nj++] = {"k_rhs_uz", vofk::k_rhs_uz, fluid->srcuz, &fluid->uz, ifrcz, ctx->d_lapx_c,
                fluid->m_dif, 2};
#endif
  jobs[nj++] = {"k_rhs_t", vofk::k_rhs_t, fluid->srct, &fluid->t, nullptr, ctx->d_lapx_c,
                fluid->t_dif, 3};
  vofk::LineJob lj{};
  lj.nlines = ny * nz;
  lj.ny = ny;
  lj.n1 = NDIMS == 3 ? ny + 2 : 0;
  lj.ca = ca; lj.cb = cb; lj.cg = cg;
  lj.images = images;
  lj.nz = nz;
  if (vof_old) {
    vofk::FusedRhs fr{};
    fr.base = base;
    fr.vof = vof_old;
    fr.curv = (const double *)interface->curv.data;
    fr.tension = interface->tension;
    fr.use_b = rkstep != 0;
    fr.keep_a = rkstep != 2;
    for (int f = 0; f < nj; f++) {
      const int kind = jobs[f].kind;
      fr.lapx[kind] = jobs[f].lapx;
      fr.dif[kind] = jobs[f].dif;
      fr.srca[kind] = (double *)jobs[f].src[0].data;
      fr.srcb[kind] = (const double *)jobs[f].src[1].data;
      fr.delta[kind] = ctx->d_delta + (size_t)f * ncell;
    }
    int xw = 0;
    const int rt = (one_kernel && images) ? pred_rows_for(ctx, &xw) : 0;
    if (rt > 0) {
      // sources, x solve and update in one kernel; the new fields go to the
      // second buffers, which then become the fields.  The buffers are
      // allocated before any batch records launches (vof_batch_create)
      if (ctx->rec && (!ctx->d_fspare[0] || !ctx->d_fspare[3]))
        return fail("one-kernel predictor: second field buffers missing in a batch");
      if (ensure_field_spares(ctx, fluid)) return 1;
      vofk::PredLines P{};
      P.rhs = fr;
      P.xw = xw;
      for (int f = 0; f < nj; f++) {
        const int kind = jobs[f].kind;
        const double * dtab = nullptr;
        if (upload_tab(ctx, kind, (int)rkstep, kind == 3 ? pt : pm, &dtab)) return 1;
        P.tab[kind] = dtab;
        P.out[kind] = ctx->d_fspare[kind];
      }
      const int R = (256 / xw) * rt;
      const size_t shm = sizeof(double) * (size_t)(NDIMS + 1) * (size_t)R * (size_t)nx;
      const dim3 gp(1u, (unsigned)((ny + R - 1) / R), (unsigned)nz);
      switch (rt) {
        case 1:
          KLB(ctx, "k_pred_lines", 256, vofk::k_pred_lines_body<1>, vofk::k_pred_lines<1>, gp, dim3(256), shm,
              ctx->grid, P);
          break;
        case 2:
          KLB(ctx, "k_pred_lines", 256, vofk::k_pred_lines_body<2>, vofk::k_pred_lines<2>, gp, dim3(256), shm,
              ctx->grid, P);
          break;
        case 4:
          KLB(ctx, "k_pred_lines", 256, vofk::k_pred_lines_body<4>, vofk::k_pred_lines<4>, gp, dim3(256), shm,
              ctx->grid, P);
          break;
        default:
          KLB(ctx, "k_pred_lines", 256, vofk::k_pred_lines_body<8>, vofk::k_pred_lines<8>, gp, dim3(256), shm,
              ctx->grid, P);
      }
      for (int f = 0; f < nj; f++) {
        const int kind = jobs[f].kind;
        void * tmp = jobs[f].field->data;
        jobs[f].field->data = ctx->d_fspare[kind];
        ctx->d_fspare[kind] = (double *)tmp;
      }
      // the last reader of the curvature of this stage is done
      if (ctx->pending_iface && ctx->iface_overlap == 1 && side_iface(ctx)) return 1;
      return 0;
    }
    const int rows = ctx->rhs_rows;
    switch (rows) {
      case 2:
        KLB(ctx, "k_rhs_march", 256, vofk::k_rhs_march_body<2>, vofk::k_rhs_march<2>, grid3(nx, (ny + 1) / 2, nz, kBlk),
            kBlk, 0, ctx->grid, fr);
        break;
      case 4:
        KLB(ctx, "k_rhs_march", 256, vofk::k_rhs_march_body<4>, vofk::k_rhs_march<4>, grid3(nx, (ny + 3) / 4, nz, kBlk),
            kBlk, 0, ctx->grid, fr);
        break;
      case 8:
        KLB(ctx, "k_rhs_march", 256, vofk::k_rhs_march_body<8>, vofk::k_rhs_march<8>, grid3(nx, (ny + 7) / 8, nz, kBlk),
            kBlk, 0, ctx->grid, fr);
        break;
      case 16:
        KLB(ctx, "k_rhs_march", 256, vofk::k_rhs_march_body<16>, vofk::k_rhs_march<16>, grid3(nx, (ny + 15) / 16, nz, kBlk),
            kBlk, 0, ctx->grid, fr);
        break;
      default:
        KLB(ctx, "k_rhs_fused", 256, vofk::k_rhs_fused_body, vofk::k_rhs_fused, grid3(nx, ny, nz, kBlk), kBlk, 0,
            ctx->grid, fr);
    }
    // the last reader of the curvature of this stage is done
    if (ctx->pending_iface && ctx->iface_overlap == 1 && side_iface(ctx)) return 1;
  }
  for (int f = 0; f < nj; f++) {
    const int kind = jobs[f].kind;
    vofk::RhsArgs a = base;
    a.ifrc = jobs[f].ifrc;
    a.lapx = jobs[f].lapx;
    a.srca = (double *)jobs[f].src[0].data;
    a.dif = jobs[f].dif;
    double * slab = ctx->d_delta + (size_t)f * ncell;
    if (vof_old) {
      // sources already formed by k_rhs_fused
    } else if (fused) {
      a.delta = slab;
      a.srcb = (const double *)jobs[f].src[1].data;
      a.srcg = nullptr;
    } else {
      a.delta = nullptr;
      a.srcb = nullptr;
      a.srcg = (double *)jobs[f].src[2].data;
    }
    const int ni = kind == 0 ? nx - 1 : nx;
    if (!vof_old) KLN(ctx, jobs[f].name, jobs[f].kern, grid3(ni, ny, nz, kBlk), kBlk, ctx->grid, a);
    const double * dtab = nullptr;
    if (upload_tab(ctx, kind, (int)rkstep, kind == 3 ? pt : pm, &dtab)) return 1;
    vofk::LineField & F = lj.f[f];
    F.tab = dtab;
    F.srca = (const double *)jobs[f].src[0].data;
    F.srcb = (const double *)jobs[f].src[1].data;
    F.srcg = (fused || vof_old) ? nullptr : (const double *)jobs[f].src[2].data;
    F.q = slab;
    F.field = (double *)jobs[f].field->data;
    F.n = ni;
    F.n0 = kind == 0 ? nx + 1 : nx + 2;
    F.i0 = 1;
  }
  {
    // lines per work-group: an LDS tile of about 16 KiB (8 lines at nx = 256,
    // the measured optimum between more resident groups and longer serial
    // sweeps per group)
    const int line_bytes = (int)(sizeof(double) * (nx | 1));
    int lpb = ctx->lines_lpb > 0 ? ctx->lines_lpb : (16384 + line_bytes - 1) / line_bytes;
    if (lpb > 64) lpb = 64;
    if (lpb < 1) lpb = 1;
    if (sizeof(double) * (nx | 1) > 65536) return fail("x solver: nx = %d exceeds the LDS line tile", nx);
    if ((size_t)lpb * (nx | 1) * sizeof(double) > 65536) lpb = (int)(65536 / (sizeof(double) * (nx | 1)));
    const size_t shm = sizeof(double) * (size_t)lpb * (nx | 1);
    const int ngroups = (lj.nlines + lpb - 1) / lpb;
    KLB(ctx, "k_lines_lds", vofk::kWThreads, vofk::k_lines_lds_body, vofk::k_lines_lds,
        dim3((unsigned)ngroups, (unsigned)nj), dim3(vofk::kWThreads), shm, lj, lpb);
  }
  if (images) return 0;
  for (int f = 0; f < nj; f++) {
    double * fld = (double *)jobs[f].field->data;
    int rc = 0;
    switch (jobs[f].kind) {
      case 0: rc = bc_ux(ctx, fld); break;
      case 1: rc = bc_uy(ctx, fld); break;
      case 2: rc = bc_uz(ctx, fld); break;
      default: rc = bc_t(ctx, fld); break;
    }
    if (rc) return 1;
  }
  return 0;
}

extern "C" int fluid_predict_field(const domain_t * domain, const size_t rkstep, const double dt,
                                   fluid_t * fluid, const interface_t * interface) {
  return predict_impl(domain, rkstep, dt, fluid, interface, 0);
}

// Poisson: both reference solvers are the same discrete problem; the FFT
// normalisation 1/(ny nz) is folded into the right-hand side as dft.c:329-330
// mixed-precision solve: FP32 transforms, FP64 eliminations, and
// `poisson_refine` FP64 residual corrections; psi always carries its halo
// images and Neumann walls (the residual reads them)
// x eliminations of the spectrum: the partitioned solve (k_thomas_part, one
// lane per segment of 2 .. 8 rows) for 4 <= nx <= 256 unless thomas_part is
// 0, else k_thomas_spec with ctx->thomas_rows rows per prefetched chunk (the
// arithmetic does not depend on it)
template <class T2>
static int launch_thomas(vof_ctx * ctx, const char * name, T2 * spec, int nlines) {
  if (ctx->thomas_part && ctx->nx >= 4 && ctx->nx <= 256) {
    // timing label of the kernel that runs
    const char * pname = sizeof(T2) == sizeof(float2) ? "k_thomas_part32" : "k_thomas_part";
    const dim3 g((unsigned)((nlines + vofk::kTpLines - 1) / vofk::kTpLines)), b(vofk::kTpLines * vofk::kTpSeg);
    KLB(ctx, pname, vofk::kTpLines * vofk::kTpSeg, (vofk::k_thomas_part_body<T2>), (vofk::k_thomas_part<T2>), g, b, 0,
        ctx->nx, nlines, ctx->nh, ctx->d_pl, ctx->d_pu, ctx->d_ey, ctx->d_ez, ctx->d_stab, spec, ctx->d_vbuf);
    return 0;
  }
  // two lanes per line (real and imaginary parts)
  const dim3 g((unsigned)((2 * (size_t)nlines + 63) / 64)), b(64);
#define TS_ARGS ctx->nx, nlines, ctx->nh, ctx->d_pl, ctx->d_pu, ctx->d_ey, ctx->d_ez, spec, ctx->d_vbuf, vbuf_store(ctx)
  switch (ctx->thomas_rows) {
    case 16:
      KLB(ctx, name, 64, (vofk::k_thomas_spec_body<T2, 16>), (vofk::k_thomas_spec<T2, 16>), g, b, 0, TS_ARGS);
      break;
    case 32:
      KLB(ctx, name, 64, (vofk::k_thomas_spec_body<T2, 32>), (vofk::k_thomas_spec<T2, 32>), g, b, 0, TS_ARGS);
      break;
    default:
      KLB(ctx, name, 64, (vofk::k_thomas_spec_body<T2, 8>), (vofk::k_thomas_spec<T2, 8>), g, b, 0, TS_ARGS);
  }
#undef TS_ARGS
  return 0;
}

static int potential_mixed(vof_ctx * ctx, size_t rkstep, double dt, fluid_t * fluid) {
  const int nx = ctx->nx, ny = ctx->ny, nz = ctx->nz;
  const double prefactor = 1. / (kRk[rkstep][2] * dt);
  const double scale = 1. / ((double)ny * (NDIMS == 3 ? (double)nz : 1.));
  const int elems = 2048;
  // 256-point transforms take the register radix-16 path, as in potential()
  const bool f256y = ny == 256 && ctx->fft_radix == 16;
  const int NC = f256y ? vofk::kF256Cols : elems / ny;
  const int nbx = (nx + 2 * NC - 1) / (2 * NC);
  const dim3 gy(nbx, NDIMS == 3 ? nz : 1), bf(vofk::kFftThreads);
#if NDIMS == 3
  const double * uz = (const double *)fluid->uz.data;
  const bool f256z = nz == 256 && ctx->fft_radix == 16;
  const int NCz = f256z ? vofk::kF256ColsZ : elems / nz;
  const dim3 gz((ctx->nh + NCz - 1) / NCz, nx);
#else
  const double * uz = nullptr;
#endif
  double * psi = (double *)fluid->psi.data;
  const int nlines = ctx->nh * nz;
  for (int pass = 0; pass <= ctx->poisson_refine; pass++) {
    // the FP64 right-hand side is kept only when a refinement pass reads it
    const int store_rhs = ctx->poisson_refine > 0;
#define RES_ARGS ctx->grid, ctx->plan_y, ctx->d_twy32, (const double *)fluid->ux.data, (const double *)fluid->uy.data, uz, \
                 prefactor, scale, ctx->d_delta, (const double *)psi, ctx->d_pl, ctx->d_pu, pass == 0 ? 0 : 1, store_rhs, \
                 ctx->d_spec32
    if (f256y) {
      KLB(ctx, "k_res_rfft32_y", vofk::kFftThreads, (vofk::k_res_rfft32_y_body<2048, true>),
          (vofk::k_res_rfft32_y<2048, true>), gy, bf, 0, RES_ARGS);
    } else {
      KLB(ctx, "k_res_rfft32_y", vofk::kFftThreads, vofk::k_res_rfft32_y_body<2048>, vofk::k_res_rfft32_y<2048>, gy, bf,
          0, RES_ARGS);
    }
#undef RES_ARGS
#if NDIMS == 3
    for (int zp = 0; zp < 2; zp++) {
      if (zp == 1 && launch_thomas<float2>(ctx, "k_thomas_spec32", ctx->d_spec32, nlines)) return 1;
      const float sign = zp == 0 ? -1.f : +1.f;
      if (f256z) {
        KLB(ctx, "k_fft32_z", vofk::kFftThreads, (vofk::k_fft_z_body<1024, float2, true>),
            (vofk::k_fft_z<1024, float2, true>), gz, bf, 0, ctx->grid, ctx->plan_z, ctx->d_twz32, ctx->d_spec32,
            ctx->nh, sign);
      } else {
        KLB(ctx, "k_fft32_z", vofk::kFftThreads, (vofk::k_fft_z_body<2048, float2>), (vofk::k_fft_z<2048, float2>), gz,
            bf, 0, ctx->grid, ctx->plan_z, ctx->d_twz32, ctx->d_spec32, ctx->nh, sign);
      }
    }
#else
    if (launch_thomas<float2>(ctx, "k_thomas_spec32", ctx->d_spec32, nlines)) return 1;
#endif
#define IRF32_ARGS ctx->grid, ctx->plan_y, ctx->d_twy32, (const float2 *)ctx->d_spec32, psi, pass == 0 ? 0 : 1
    if (f256y) {
      KLB(ctx, "k_irfft32_y_psi", vofk::kFftThreads, (vofk::k_irfft32_y_psi_body<2048, true>),
          (vofk::k_irfft32_y_psi<2048, true>), gy, bf, 0, IRF32_ARGS);
    } else {
      KLB(ctx, "k_irfft32_y_psi", vofk::kFftThreads, vofk::k_irfft32_y_psi_body<2048>, vofk::k_irfft32_y_psi<2048>, gy,
          bf, 0, IRF32_ARGS);
    }
#undef IRF32_ARGS
  }
  return 0;
}

static int potential(const domain_t * domain, size_t rkstep, double dt, fluid_t * fluid, int images) {
  CTX_OR_FAIL(domain);
  if (rkstep > 2) return fail("rkstep %zu out of range", rkstep);
  if (ctx->poisson_refine >= 0) return potential_mixed(ctx, rkstep, dt, fluid);
  const int nx = ctx->nx, ny = ctx->ny, nz = ctx->nz;
  const double norm = (double)ny * (NDIMS == 3 ? (double)nz : 1.);
  const double prefactor = 1. / (kRk[rkstep][2] * dt) / norm;
  const int elems = (ctx->fft_elems == 1024 && ny <= 1024 && (NDIMS == 2 || nz <= 1024)) ? 1024 : 2048;
  // 256-point transforms take the register radix-16 path (8 columns per
  // work-group in y, 16 in z), other lengths the mixed-radix Stockham path
  const bool f256y = ny == 256 && ctx->fft_radix == 16;
  const int NC = f256y ? vofk::kF256Cols : elems / ny;
  const int nbx = (nx + 2 * NC - 1) / (2 * NC);
  const dim3 gy(nbx, NDIMS == 3 ? nz : 1), bf(vofk::kFftThreads);
#if NDIMS == 3
  const double * uz = (const double *)fluid->uz.data;
#else
  const double * uz = nullptr;
#endif
#define DIV_ARGS ctx->grid, ctx->plan_y, ctx->d_twy, (const double *)fluid->ux.data, (const double *)fluid->uy.data, uz, \
                 prefactor, ctx->d_spec
  if (f256y) {
    KLB(ctx, "k_div_rfft_y", vofk::kFftThreads, (vofk::k_div_rfft_y_body<1024, true>), (vofk::k_div_rfft_y<1024, true>), gy,
        bf, 0, DIV_ARGS);
  } else if (elems == 1024) {
    KLB(ctx, "k_div_rfft_y", vofk::kFftThreads, vofk::k_div_rfft_y_body<1024>, vofk::k_div_rfft_y<1024>, gy, bf, 0, DIV_ARGS);
  } else {
    KLB(ctx, "k_div_rfft_y", vofk::kFftThreads, vofk::k_div_rfft_y_body<2048>, vofk::k_div_rfft_y<2048>, gy, bf, 0, DIV_ARGS);
  }
#undef DIV_ARGS
#if NDIMS == 3
  const bool f256z = nz == 256 && ctx->fft_radix == 16;
  const int NCz = f256z ? vofk::kF256ColsZ : elems / nz;
  const dim3 gz((ctx->nh + NCz - 1) / NCz, nx);
  for (int pass = 0; pass < 2; pass++) {
    if (pass == 1) {
      if (ctx->pending_iface && ctx->iface_overlap == 2 && side_iface(ctx)) return 1;
      if (launch_thomas<double2>(ctx, "k_thomas_spec", ctx->d_spec, ctx->nh * nz)) return 1;
    }
    const double sign = pass == 0 ? -1. : +1.;
    if (f256z) {
      KLB(ctx, "k_fft_z", vofk::kFftThreads, (vofk::k_fft_z_body<1024, double2, true>), (vofk::k_fft_z<1024, double2, true>),
          gz, bf, 0, ctx->grid, ctx->plan_z, ctx->d_twz, ctx->d_spec, ctx->nh, sign);
    } else if (elems == 1024) {
      KLB(ctx, "k_fft_z", vofk::kFftThreads, (vofk::k_fft_z_body<1024, double2>), (vofk::k_fft_z<1024, double2>), gz, bf, 0,
          ctx->grid, ctx->plan_z, ctx->d_twz, ctx->d_spec, ctx->nh, sign);
    } else {
      KLB(ctx, "k_fft_z", vofk::kFftThreads, (vofk::k_fft_z_body<2048, double2>), (vofk::k_fft_z<2048, double2>), gz, bf, 0,
          ctx->grid, ctx->plan_z, ctx->d_twz, ctx->d_spec, ctx->nh, sign);
    }
  }
#else
  {
    if (ctx->pending_iface && ctx->iface_overlap == 2 && side_iface(ctx)) return 1;
    if (launch_thomas<double2>(ctx, "k_thomas_spec", ctx->d_spec, ctx->nh * nz)) return 1;
  }
#endif
#define IRF_ARGS ctx->grid, ctx->plan_y, ctx->d_twy, ctx->d_spec, (double *)fluid->psi.data, images
  if (f256y) {
    KLB(ctx, "k_irfft_y_psi", vofk::kFftThreads, (vofk::k_irfft_y_psi_body<1024, true>), (vofk::k_irfft_y_psi<1024, true>),
        gy, bf, 0, IRF_ARGS);
  } else if (elems == 1024) {
    KLB(ctx, "k_irfft_y_psi", vofk::kFftThreads, vofk::k_irfft_y_psi_body<1024>, vofk::k_irfft_y_psi<1024>, gy, bf, 0,
        IRF_ARGS);
  } else {
    KLB(ctx, "k_irfft_y_psi", vofk::kFftThreads, vofk::k_irfft_y_psi_body<2048>, vofk::k_irfft_y_psi<2048>, gy, bf, 0,
        IRF_ARGS);
  }
#undef IRF_ARGS
  return images ? 0 : bc_neumann(ctx, (double *)fluid->psi.data);
}

extern "C" int fluid_compute_potential_dft(const domain_t * domain, const size_t rkstep,
                                           const double dt, fluid_t * fluid) {
  return potential(domain, rkstep, dt, fluid, 0);
}

extern "C" int fluid_compute_potential_dct(const domain_t * domain, const size_t rkstep,
                                           const double dt, fluid_t * fluid) {
  return potential(domain, rkstep, dt, fluid, 0);
}

// correct_velocity/{main,ux,uy}.c
extern "C" int fluid_correct_velocity(const domain_t * domain, const size_t rkstep,
                                      const double dt, fluid_t * fluid) {
  CTX_OR_FAIL(domain);
  if (rkstep > 2) return fail("rkstep %zu out of range", rkstep);
  const double prefactor = kRk[rkstep][2] * dt;
#if NDIMS == 3
  double * uz = (double *)fluid->uz.data;
#else
  double * uz = nullptr;
#endif
  KLN(ctx, "k_correct", vofk::k_correct, grid3(ctx->nx, ctx->ny, ctx->nz, kBlk), kBlk,
                     ctx->grid, prefactor, (const double *)fluid->psi.data,
                     (double *)fluid->ux.data, (double *)fluid->uy.data, uz);
  LAUNCHCHK();
  if (bc_ux(ctx, (double *)fluid->ux.data)) return 1;
  if (bc_uy(ctx, (double *)fluid->uy.data)) return 1;
#if NDIMS == 3
  if (bc_uz(ctx, uz)) return 1;
#endif
  return 0;
}

// update_pressure.c:87-109
extern "C" int fluid_update_pressure(const domain_t * domain, const size_t rkstep, const double dt,
                                     fluid_t * fluid) {
  CTX_OR_FAIL(domain);
  if (rkstep > 2) return fail("rkstep %zu out of range", rkstep);
  const double prefactor = 0.5 * kRk[rkstep][2] * dt * fluid->m_dif;
  KLN(ctx, "k_pressure", vofk::k_pressure, grid3(ctx->nx, ctx->ny, ctx->nz, kBlk), kBlk,
                     ctx->grid, prefactor, (const double *)fluid->psi.data, (double *)fluid->p.data);
  LAUNCHCHK();
  return bc_neumann(ctx, (double *)fluid->p.data);
}

// ---------------------------------------------------------------------------
// interface

extern "C" int interface_init(const char dirname_ic[], const domain_t * domain, interface_t * interface) {
  CTX_OR_FAIL(domain);
  if (array_prepare(domain, "vof", sizeof(double), &interface->vof)) return 1;
  if (array_prepare(domain, "ifrcx", sizeof(double), &interface->ifrcx)) return 1;
  if (array_prepare(domain, "ifrcy", sizeof(double), &interface->ifrcy)) return 1;
#if NDIMS == 3
  if (array_prepare(domain, "ifrcz", sizeof(double), &interface->ifrcz)) return 1;
#endif
  if (array_prepare(domain, "dvof", sizeof(vector_t), &interface->dvof)) return 1;
  if (array_prepare(domain, "normal", sizeof(normal_t), &interface->normal)) return 1;
  if (array_prepare(domain, "curv", sizeof(double), &interface->curv)) return 1;
  if (array_prepare(domain, "flxx", sizeof(double), &interface->flxx)) return 1;
  if (array_prepare(domain, "flxy", sizeof(double), &interface->flxy)) return 1;
#if NDIMS == 3
  if (array_prepare(domain, "flxz", sizeof(double), &interface->flxz)) return 1;
#endif
  for (int m = 0; m < 2; m++)
    if (array_prepare(domain, "src", sizeof(double), &interface->src[m])) return 1;
  if (array_load(domain, dirname_ic, "vof", &interface->vof)) return 1;
  if (interface_update_boundaries_vof(domain, &interface->vof)) return 1;
  double We = 0.;
  if (param_get(ctx, P_WE, &We)) return 1;
  interface->tension = 1. / We;
  HIPCHK(hipStreamSynchronize(ctx->stream));
  return 0;
}

extern "C" int interface_finalise(const domain_t * domain, interface_t * interface) {
  if (ctx_of(domain)) (void)hipStreamSynchronize(ctx_of(domain)->stream);
  array_free(&interface->vof);
  array_free(&interface->ifrcx);
  array_free(&interface->ifrcy);
#if NDIMS == 3
  array_free(&interface->ifrcz);
  array_free(&interface->flxz);
#endif
  array_free(&interface->dvof);
  array_free(&interface->normal);
  array_free(&interface->curv);
  array_free(&interface->flxx);
  array_free(&interface->flxy);
  array_free(&interface->src[0]);
  array_free(&interface->src[1]);
  return 0;
}

extern "C" int interface_save(const char dirname[], const domain_t * domain, const interface_t * interface) {
  CTX_OR_FAIL(domain);
  (void)ctx;
  if (npy::write(dirname, "surface_tension", kNpyDouble, {}, &interface->tension, sizeof(double))) return 1;
  return array_dump(domain, dirname, "vof", &interface->vof);
}

extern "C" int interface_update_boundaries_vof(const domain_t * domain, array_t * vof) {
  CTX_OR_FAIL(domain);
  return bc_vof(ctx, (double *)vof->data);
}

extern "C" int interface_compute_curvature_tensor(const domain_t * domain, interface_t * interface) {
  CTX_OR_FAIL(domain);
  const int nx = ctx->nx, ny = ctx->ny, nz = ctx->nz;
  KLN(ctx, "k_gradient", vofk::k_gradient, grid3(nx + 1, ny + 3, nz + 3, kBlk), kBlk,
                     ctx->grid, (const double *)interface->vof.data, (double *)interface->dvof.data);
  LAUNCHCHK();
  KLN(ctx, "k_normal_curv", vofk::k_normal_curv, grid3(nx, ny + 2, nz + 2, kBlk), kBlk,
                     ctx->grid, (const double *)interface->vof.data,
                     (const double *)interface->dvof.data, (double *)interface->normal.data,
                     (double *)interface->curv.data);
  LAUNCHCHK();
  return 0;
}

extern "C" int interface_compute_force(const domain_t * domain, interface_t * interface) {
  CTX_OR_FAIL(domain);
#if NDIMS == 3
  double * ifrcz = (double *)interface->ifrcz.data;
#else
  double * ifrcz = nullptr;
#endif
  KLN(ctx, "k_force", vofk::k_force, grid3(ctx->nx, ctx->ny, ctx->nz, kBlk), kBlk,
                     ctx->grid, interface->tension, (const double *)interface->vof.data,
                     (const double *)interface->curv.data, (double *)interface->ifrcx.data,
                     (double *)interface->ifrcy.data, ifrcz);
  LAUNCHCHK();
  return 0;
}

extern "C" int interface_update_vof(const domain_t * domain, const size_t rkstep, const double dt,
                                    const fluid_t * fluid, interface_t * interface) {
  CTX_OR_FAIL(domain);
  if (rkstep > 2) return fail("rkstep %zu out of range", rkstep);
  // reset_srcs (update/main.c:28-45): swap; src_a is fully rewritten below
  if (rkstep != 0) swap_data(&interface->src[0], &interface->src[1]);
  const int nx = ctx->nx, ny = ctx->ny, nz = ctx->nz;
#if NDIMS == 3
  const double * uz = (const double *)fluid->uz.data;
  double * flxz = (double *)interface->flxz.data;
#else
  const double * uz = nullptr;
  double * flxz = nullptr;
#endif
  KLN(ctx, "k_flux", vofk::k_flux, grid3(nx, ny + 1, nz + 1, kBlk), kBlk, ctx->grid,
                     (const double *)fluid->ux.data, (const double *)fluid->uy.data, uz,
                     (const double *)interface->vof.data, (const double *)interface->normal.data,
                     (double *)interface->flxx.data, (double *)interface->flxy.data, flxz);
  LAUNCHCHK();
  const double dta = dt * kRk[rkstep][0];
  const double dtb = dt * kRk[rkstep][1];
  KLN(ctx, "k_vof_update", vofk::k_vof_update, grid3(nx, ny, nz, kBlk), kBlk, ctx->grid,
                     (const double *)interface->flxx.data, (const double *)interface->flxy.data,
                     (const double *)flxz, (double *)interface->src[0].data,
                     (const double *)interface->src[1].data, (double *)interface->vof.data, dta,
                     dtb, (int)(rkstep != 0));
  LAUNCHCHK();
  return bc_vof(ctx, (double *)interface->vof.data);
}

// ---------------------------------------------------------------------------
// integrate (src/integrate.c:11-68)

enum { F_IFACE = 1, F_VOF = 2, F_RHS = 4, F_LINES = 8, F_PSI = 16, F_CORR = 32, F_PRED = 64 };

// the VOF ping-pong buffer of the fused VOF update
static int ensure_vof_spare(vof_ctx * ctx, const interface_t * interface) {
  if (ctx->d_vof_spare) return 0;
  ctx->vof_bytes = interface->vof.datasize;
  HIPCHK(hipMalloc(&ctx->d_vof_spare, ctx->vof_bytes));
  HIPCHK(hipMemsetAsync(ctx->d_vof_spare, 0, ctx->vof_bytes, ctx->stream));
  return 0;
}

static int integrate_fused_stages(const domain_t * domain, fluid_t * fluid, interface_t * interface,
                                  const double dtv, size_t rk0 = 0, size_t rk1 = 3);

static int launch_iface(vof_ctx * ctx, interface_t * interface);

// the fused schedule may use the side stream (not while launches are
// recorded for a batch, the adjoint records stage states, or per-kernel
// timing is on)
static bool overlap_ok(const vof_ctx * ctx) {
  const int fz = ctx->fuse;
  return ctx->iface_overlap && (fz & F_IFACE) && (fz & F_VOF) && (fz & F_RHS) && !ctx->rec && !ctx->tape &&
         !ctx->prof_on && !ctx->freeze_mode;
}

// the side stream and its events, created on first use (a context that
// never overlaps keeps one stream: with few hardware queues, idle extra
// streams still take queue slots from the main streams of other contexts)
static int ensure_side(vof_ctx * ctx) {
  if (ctx->side) return 0;
  HIPCHK(hipEventCreateWithFlags(&ctx->ev_rhs, hipEventDisableTiming));
  HIPCHK(hipEventCreateWithFlags(&ctx->ev_iface, hipEventDisableTiming));
  HIPCHK(hipEventCreateWithFlags(&ctx->ev_main, hipEventDisableTiming));
  HIPCHK(hipStreamCreate(&ctx->side));
  return 0;
}

static int integrate_fused(const domain_t * domain, fluid_t * fluid, interface_t * interface, double * dt) {
  vof_ctx * ctx = ctx_of(domain);
  if ((ctx->fuse & F_VOF) && ensure_vof_spare(ctx, interface)) return 1;
  if ((ctx->fuse & F_PRED) && ensure_field_spares(ctx, fluid)) return 1;
  if (overlap_ok(ctx)) {
    if (ensure_side(ctx)) return 1;
    // the first stage's interface kernel needs no dt: it runs on the main
    // stream while dt is reduced on the side stream (after everything
    // issued before on the main stream, which may still write the velocity)
    HIPCHK(hipEventRecord(ctx->ev_main, ctx->stream));
    HIPCHK(hipStreamWaitEvent(ctx->side, ctx->ev_main, 0));
    if (launch_iface(ctx, interface)) return 1;
    hipStream_t main_stream = ctx->stream;
    ctx->stream = ctx->side;
    const int rc = fluid_decide_dt(domain, fluid, dt);
    ctx->stream = main_stream;
    if (rc) return 1;
    ctx->iface_ready = true;
    const int rs = integrate_fused_stages(domain, fluid, interface, *dt);
    ctx->iface_ready = false;
    return rs;
  }
  if (fluid_decide_dt(domain, fluid, dt)) return 1;
  return integrate_fused_stages(domain, fluid, interface, *dt);
}

#if NDIMS == 2
// the decision slot of the current RK stage (allocated while recording)
static int freeze_slot(vof_ctx * ctx, vofk::Freeze * fz) {
  const size_t nx = (size_t)ctx->nx, ny = (size_t)ctx->ny;
  const size_t nc = (nx + 1) * (ny + 3), ncell = (nx + 2) * (ny + 2), nfx = (nx + 1) * (ny + 2);
  const size_t flags = (nc + ncell + nfx + ncell + 7) / 8 * 8;
  if (ctx->freeze_mode == 1 && ctx->freeze_pos == ctx->freeze_slots.size()) {
    void * b = nullptr;
    HIPCHK(hipMalloc(&b, flags + sizeof(double) * 2 * nc));
    ctx->freeze_slots.push_back(b);
  }
  if (ctx->freeze_mode == 2 && ctx->freeze_pos >= ctx->freeze_n)
    return fail("vof_set_freeze: replay past the %zu recorded RK stages", ctx->freeze_n);
  unsigned char * b = (unsigned char *)ctx->freeze_slots[ctx->freeze_pos];
  fz->corner = b;
  fz->cell = b + nc;
  fz->fx = b + nc + ncell;
  fz->fy = b + nc + ncell + nfx;
  fz->dvof = (double *)(b + flags);
  fz->floor = ctx->adj_floor;
  return 0;
}
#endif

// RK stages rk0 .. rk1-1 (all three by default) of one step with dt known
// fused curvature tensor (curvature_tensor.c:178-186) of the current vof
// on ctx->stream
static int launch_iface(vof_ctx * ctx, interface_t * interface) {
#if NDIMS == 2
  if (ctx->freeze_mode) {
    vofk::Freeze fz;
    if (freeze_slot(ctx, &fz)) return 1;
    const dim3 gi((unsigned)((ctx->nx + vofk::kIX - 1) / vofk::kIX), (unsigned)((ctx->ny + vofk::kIY - 1) / vofk::kIY), 1);
    const double * v = (const double *)interface->vof.data;
    if (ctx->freeze_mode == 1)
      KLN(ctx, "k_iface_fused", vofk::k_iface_fused_fz<1>, gi, dim3(vofk::kIThreads), ctx->grid, v,
          (double *)interface->normal.data, (double *)interface->curv.data, fz);
    else
      KLN(ctx, "k_iface_fused", vofk::k_iface_fused_fz<2>, gi, dim3(vofk::kIThreads), ctx->grid, v,
          (double *)interface->normal.data, (double *)interface->curv.data, fz);
    return 0;
  }
#endif
  const int nx = ctx->nx, ny = ctx->ny, nz = ctx->nz;
  const double * v = (const double *)interface->vof.data;
  double * nrm = (double *)interface->normal.data;
  double * cu = (double *)interface->curv.data;
#if NDIMS == 3
  const int kz = ctx->iface_planes;
  const dim3 gm((unsigned)((nx + vofk::kMX - 1) / vofk::kMX), (unsigned)((ny + vofk::kMY - 1) / vofk::kMY),
                (unsigned)((nz + (kz > 0 ? kz : 1) - 1) / (kz > 0 ? kz : 1)));
  const dim3 bm(vofk::kMThreads);
  if (kz == 4) {
    KLB(ctx, "k_iface_march", vofk::kMThreads, vofk::k_iface_march_body<4>, vofk::k_iface_march<4>, gm, bm, 0,
        ctx->grid, v, nrm, cu);
  } else if (kz == 8) {
    KLB(ctx, "k_iface_march", vofk::kMThreads, vofk::k_iface_march_body<8>, vofk::k_iface_march<8>, gm, bm, 0,
        ctx->grid, v, nrm, cu);
  } else if (kz == 16) {
    KLB(ctx, "k_iface_march", vofk::kMThreads, vofk::k_iface_march_body<16>, vofk::k_iface_march<16>, gm, bm, 0,
        ctx->grid, v, nrm, cu);
  } else
#endif
  {
    const dim3 gi((unsigned)((nx + vofk::kIX - 1) / vofk::kIX), (unsigned)((ny + vofk::kIY - 1) / vofk::kIY),
                  (unsigned)((nz + vofk::kIZ - 1) / vofk::kIZ));
    KLB(ctx, "k_iface_fused", vofk::kIThreads, vofk::k_iface_fused_body, vofk::k_iface_fused, gi,
        dim3(vofk::kIThreads), 0, ctx->grid, v, nrm, cu);
  }
  return 0;
}

static int integrate_fused_stages(const domain_t * domain, fluid_t * fluid, interface_t * interface,
                                  const double dtv, size_t rk0, size_t rk1) {
  vof_ctx * ctx = ctx_of(domain);
  const int fz = ctx->fuse;
  const int nx = ctx->nx, ny = ctx->ny, nz = ctx->nz;
  // the field buffers are read at each use: the one-kernel predictor swaps
  // ux, uy, uz, t with their second buffers every stage
#if NDIMS == 3
  auto uzp = [&]() { return (double *)fluid->uz.data; };
#else
  auto uzp = [&]() { return (double *)nullptr; };
#endif
  // the interface kernel of stage rk + 1 reads only the vof that stage rk's
  // VOF update wrote, and writes the normals and curvature whose last reader
  // in stage rk is the source kernel: it runs on the side stream from the end
  // of that kernel, beside the x solve, the Poisson solve and the
  // correction (not while launches are recorded for a batch, the adjoint
  // records stage states, or per-kernel timing is on)
  const bool overlap = overlap_ok(ctx);
  if (overlap && ensure_side(ctx)) return 1;
  for (size_t rk = rk0; rk < rk1; rk++) {
    if (ctx->tape) {
      // the adjoint's recomputation records each stage's input state
      double ** t = ctx->tape + 3 * rk;
      HIPCHK(hipMemcpyAsync(t[0], interface->vof.data, interface->vof.datasize, hipMemcpyDeviceToDevice,
                            ctx->stream));
      HIPCHK(hipMemcpyAsync(t[1], fluid->ux.data, fluid->ux.datasize, hipMemcpyDeviceToDevice, ctx->stream));
      HIPCHK(hipMemcpyAsync(t[2], fluid->uy.data, fluid->uy.datasize, hipMemcpyDeviceToDevice, ctx->stream));
    }
    // curvature tensor (curvature_tensor.c:178-186)
    if (fz & F_IFACE) {
      if (rk != rk0 && overlap) {
        // launched on the side stream during the previous stage
        HIPCHK(hipStreamWaitEvent(ctx->stream, ctx->ev_iface, 0));
      } else if (rk == rk0 && ctx->iface_ready) {
        // launched by integrate before dt was reduced
        ctx->iface_ready = false;
      } else if (launch_iface(ctx, interface)) {
        return 1;
      }
    } else {
      if (interface_compute_curvature_tensor(domain, interface)) return 1;
    }
    const bool rhs_fused = (fz & F_RHS) && (fz & F_VOF);
    if (!rhs_fused && interface_compute_force(domain, interface)) return 1;
    // VOF update (update/main.c:112-126)
    const double * vof_old = nullptr;
    if (fz & F_VOF) {
      if (rk != 0) swap_data(&interface->src[0], &interface->src[1]);
      const double dta = dtv * kRk[rk][0];
      const double dtb = dtv * kRk[rk][1];
#if NDIMS == 2
      if (ctx->freeze_mode) {
        vofk::Freeze fzs;
        if (freeze_slot(ctx, &fzs)) return 1;
        if (ctx->freeze_mode == 1)
          KLN(ctx, "k_vof_fused", vofk::k_vof_fused_fz<1>, grid3(nx, ny, nz, kBlk), kBlk, ctx->grid,
              (const double *)fluid->ux.data, (const double *)fluid->uy.data, (const double *)interface->vof.data,
              (const double *)interface->normal.data, (double *)interface->src[0].data,
              (const double *)interface->src[1].data, ctx->d_vof_spare, dta, dtb, (int)(rk != 0), (int)(rk != 2),
              fzs);
        else
          KLN(ctx, "k_vof_fused", vofk::k_vof_fused_fz<2>, grid3(nx, ny, nz, kBlk), kBlk, ctx->grid,
              (const double *)fluid->ux.data, (const double *)fluid->uy.data, (const double *)interface->vof.data,
              (const double *)interface->normal.data, (double *)interface->src[0].data,
              (const double *)interface->src[1].data, ctx->d_vof_spare, dta, dtb, (int)(rk != 0), (int)(rk != 2),
              fzs);
        ctx->freeze_pos++;
        if (ctx->freeze_mode == 1) ctx->freeze_n = ctx->freeze_pos;
      } else
#endif
      KLB(ctx, "k_vof_fused", 256, vofk::k_vof_fused_body, vofk::k_vof_fused, grid3(nx, ny, nz, kBlk), kBlk, 0,
          ctx->grid, (const double *)fluid->ux.data, (const double *)fluid->uy.data, (const double *)uzp(),
          (const double *)interface->vof.data, (const double *)interface->normal.data,
          (double *)interface->src[0].data, (const double *)interface->src[1].data,
          ctx->d_vof_spare, dta, dtb, (int)(rk != 0), (int)(rk != 2));
      void * tmp = interface->vof.data;
      interface->vof.data = ctx->d_vof_spare;
      ctx->d_vof_spare = (double *)tmp;
      vof_old = ctx->d_vof_spare;
    } else {
      if (interface_update_vof(domain, rk, dtv, fluid, interface)) return 1;
    }
    // predictor (predict/main.c:62-79); the next stage's interface kernel
    // goes to the side stream after the source kernel (iface_overlap 1) or
    // before the Thomas pass of the Poisson solve (2)
    ctx->pending_iface = (overlap && rk + 1 < rk1) ? interface : nullptr;
    if (predict_impl(domain, rk, dtv, fluid, interface, 1, rhs_fused ? vof_old : nullptr,
                     (fz & F_LINES) ? 1 : 0, (fz & F_PRED) != 0)) {
      ctx->pending_iface = nullptr;
      return 1;
    }
    // Poisson (both reference solvers are the same discrete problem)
    if (potential(domain, rk, dtv, fluid, (fz & F_PSI) ? 1 : 0)) {
      ctx->pending_iface = nullptr;
      return 1;
    }
    if (ctx->pending_iface && side_iface(ctx)) return 1;
    // correction and pressure (correct_velocity/main.c:15-27, update_pressure.c:87-109)
    if (fz & F_CORR) {
      const double pref_c = kRk[rk][2] * dtv;
      const double pref_p = 0.5 * kRk[rk][2] * dtv * fluid->m_dif;
      KLB(ctx, "k_corr_press", 256, vofk::k_corr_press_body, vofk::k_corr_press, grid3(nx, ny, nz, kBlk), kBlk, 0,
          ctx->grid, pref_c, pref_p, (const double *)fluid->psi.data, (double *)fluid->ux.data,
          (double *)fluid->uy.data, uzp(), (double *)fluid->p.data);
    } else {
      if (fluid_correct_velocity(domain, rk, dtv, fluid)) return 1;
      if (fluid_update_pressure(domain, rk, dtv, fluid)) return 1;
    }
  }
  return 0;
}

extern "C" int integrate(const domain_t * domain, fluid_t * fluid, interface_t * interface, double * dt) {
  CTX_OR_FAIL(domain);
  if (ctx->schedule == 1) return integrate_fused(domain, fluid, interface, dt);
  bool uniform = false;
  domain_check_x_grid_is_uniform(domain, &uniform);
  if (fluid_decide_dt(domain, fluid, dt)) return 1;
  for (size_t rk = 0; rk < 3; rk++) {
    if (interface_compute_curvature_tensor(domain, interface)) return 1;
    if (interface_compute_force(domain, interface)) return 1;
    if (interface_update_vof(domain, rk, *dt, fluid, interface)) return 1;
    if (fluid_predict_field(domain, rk, *dt, fluid, interface)) return 1;
    if (uniform) {
      if (fluid_compute_potential_dct(domain, rk, *dt, fluid)) return 1;
    } else {
      if (fluid_compute_potential_dft(domain, rk, *dt, fluid)) return 1;
    }
    if (fluid_correct_velocity(domain, rk, *dt, fluid)) return 1;
    if (fluid_update_pressure(domain, rk, *dt, fluid)) return 1;
  }
  return 0;
}

// tile parameters: "fft_elems" (1024 or 2048 complex per LDS buffer),
// "lines_lpb" (lines per work-group of the x solver, 0 = automatic),
// "xcd_bands" (band tile order of the 3D stencil kernels), "rhs_rows"
// (rows of one column marched per thread by the source kernel: 1, 2, 4, 8,
// 16), "pred_rows" (rows marched per thread by the one-kernel predictor: 1,
// 2, 4, 8; fewer where the work-group's x lines would not fit 64 KiB of LDS)
// and "iface_planes" (z planes marched per work-group by the 3D interface
// kernel: 0 = 32 x 4 x 4 blocks, 4, 8, 16)
extern "C" int vof_set_tuning(domain_t * domain, const char name[], int value) {
  CTX_OR_FAIL(domain);
  if (!strcmp(name, "fft_elems")) {
    if (value != 1024 && value != 2048) return fail("fft_elems must be 1024 or 2048");
    if (value < ctx->ny || (NDIMS == 3 && value < ctx->nz)) return fail("fft_elems below the FFT length");
    ctx->fft_elems = value;
    return 0;
  }
  if (!strcmp(name, "iface_overlap")) {
    if (value < 0 || value > 2) return fail("iface_overlap must be 0, 1 or 2");
    ctx->iface_overlap = value;
    return 0;
  }
  if (!strcmp(name, "thomas_rows")) {
    if (value != 8 && value != 16 && value != 32) return fail("thomas_rows must be 8, 16 or 32");
    ctx->thomas_rows = value;
    return 0;
  }
  if (!strcmp(name, "thomas_part")) {
    if (value != 0 && value != 1) return fail("thomas_part must be 0 or 1");
    ctx->thomas_part = value;
    return 0;
  }
  if (!strcmp(name, "fft_radix")) {
    if (value != 4 && value != 16) return fail("fft_radix must be 4 or 16");
    ctx->fft_radix = value;
    return 0;
  }
  if (!strcmp(name, "xcd_bands")) {
    ctx->grid.bands = value ? 1 : 0;
    return 0;
  }
  if (!strcmp(name, "adj_floor_exp")) {
    if (value < 0 || value > 16) return fail("adj_floor_exp must be in 0..16");
    ctx->adj_floor = pow(10., -(double)value);
    return 0;
  }
  if (!strcmp(name, "lines_lpb")) {
    if (value < 0 || value > 64) return fail("lines_lpb must be in 0..64");
    ctx->lines_lpb = value;
    return 0;
  }
  if (!strcmp(name, "iface_planes")) {
    if (value != 0 && value != 4 && value != 8 && value != 16) return fail("iface_planes must be 0, 4, 8 or 16");
    ctx->iface_planes = value;
    return 0;
  }
  if (!strcmp(name, "rhs_rows")) {
    if (value != 1 && value != 2 && value != 4 && value != 8 && value != 16) return fail("rhs_rows must be 1, 2, 4, 8 or 16");
    ctx->rhs_rows = value;
    return 0;
  }
  if (!strcmp(name, "pred_rows")) {
    if (value != 1 && value != 2 && value != 4 && value != 8) return fail("pred_rows must be 1, 2, 4 or 8");
    ctx->pred_rows = value;
    return 0;
  }
  return fail("vof_set_tuning: unknown parameter %s", name);
}

// Poisson precision: refinements < 0 selects the plain FP64 solver; >= 0 the
// FP32-transform solver followed by that many FP64 residual corrections
extern "C" int vof_set_poisson_mixed(domain_t * domain, int refinements) {
  CTX_OR_FAIL(domain);
  if (refinements > 16) return fail("vof_set_poisson_mixed: at most 16 refinements");
  if (refinements >= 0 && !ctx->d_spec32) {
    if (ctx->ny > 2048 || (NDIMS == 3 && ctx->nz > 2048)) return fail("mixed Poisson: FFT length above 2048");
    const size_t nline = (size_t)ctx->nh * ctx->nz;
    HIPCHK(hipMalloc(&ctx->d_spec32, sizeof(float2) * ctx->nx * nline));
    for (int d = 0; d < (NDIMS == 3 ? 2 : 1); d++) {
      const int n = d == 0 ? ctx->ny : ctx->nz;
      std::vector<double2> tw = make_twiddles(n);
      std::vector<float2> tw32(n);
      for (int m = 0; m < n; m++) tw32[m] = make_float2((float)tw[m].x, (float)tw[m].y);
      float2 ** dst = d == 0 ? &ctx->d_twy32 : &ctx->d_twz32;
      HIPCHK(hipMalloc(dst, sizeof(float2) * n));
      HIPCHK(hipMemcpy(*dst, tw32.data(), sizeof(float2) * n, hipMemcpyHostToDevice));
    }
  }
  ctx->poisson_refine = refinements < 0 ? -1 : refinements;
  return 0;
}

// ---------------------------------------------------------------------------
// batch of cases advanced together (see "Batched execution" above)

__global__ void k_gather_dt(const double * const * __restrict__ src, int n, double * __restrict__ dst) {
  const int c = (int)(blockIdx.x * blockDim.x + threadIdx.x);
  if (c < n) dst[c] = *src[c];
}

static void batch_drop_graph(vof_batch * B, int ph) {
  if (B->gexec[ph]) (void)hipGraphExecDestroy(B->gexec[ph]);
  B->gexec[ph] = nullptr;
  B->gsig[ph].clear();
}

static int batch_stage_reserve(vof_batch * B, int ph, size_t bytes) {
  if (bytes <= B->stage_bytes[ph]) return 0;
  HIPCHK(hipStreamSynchronize(B->stream));
  batch_drop_graph(B, ph);
  host_free(B->h_stage[ph]);
  dev_free(B->d_stage[ph]);
  B->h_stage[ph] = nullptr;
  B->d_stage[ph] = nullptr;
  B->stage_bytes[ph] = 0;
  const size_t nb = bytes + bytes / 2 + 4096;
  HIPCHK(hipHostMalloc(&B->h_stage[ph], nb));
  HIPCHK(hipMalloc(&B->d_stage[ph], nb));
  B->stage_bytes[ph] = nb;
  return 0;
}

// merge the recorded launches of every case, launch by launch, on B->stream.
// Cases whose i-th launch is the same kernel (instantiation and work-group
// shape) share one merged launch; sizes that select another instantiation
// (e.g. the 256-point transforms) form a group of their own.
struct BatchGroup {
  size_t r;                  // launch number
  std::vector<int> cases;
};

// phase ph of the step: `pre` and `post` enqueue the phase's other copies
// and launches (Thomas tables before a stage; the dt gather after phase 0).
// With graphs on, the whole phase is captured once and replayed while its
// signature (groups, kernel instantiations, grids, table offsets, staging
// buffers) is unchanged; the argument tables themselves are refilled on the
// host every step and copied by the graph's first node.
static int batch_run(vof_batch * B, int ph, const std::function<int()> & pre, const std::function<int()> & post) {
  const size_t n = B->ctx.size();
  const size_t nrec = B->rec[0].recs.size();
  for (size_t c = 1; c < n; c++) {
    if (B->rec[c].recs.size() != nrec) return fail("batch: case %zu records %zu launches, case 0 %zu", c,
                                                   B->rec[c].recs.size(), nrec);
  }
  std::vector<BatchGroup> groups;
  for (size_t r = 0; r < nrec; r++) {
    const size_t g0 = groups.size();
    for (size_t c = 0; c < n; c++) {
      const LaunchRec & rc = B->rec[c].recs[r];
      size_t gi = g0;
      for (; gi < groups.size(); gi++) {
        const LaunchRec & rg = B->rec[groups[gi].cases[0]].recs[r];
        if (rg.merge == rc.merge && rg.block.x == rc.block.x && rg.block.y == rc.block.y && rg.block.z == rc.block.z)
          break;
      }
      if (gi == groups.size()) groups.push_back({r, {}});
      groups[gi].cases.push_back((int)c);
    }
  }
  const size_t ng = groups.size();
  std::vector<size_t> off(ng + 1, 0);
  for (size_t q = 0; q < ng; q++) {
    const LaunchRec & r0 = B->rec[groups[q].cases[0]].recs[groups[q].r];
    const size_t bytes = r0.merge(B, groups[q].r, groups[q].cases, nullptr, nullptr);
    off[q + 1] = off[q] + ((bytes + 255) & ~(size_t)255);
  }
  // the previous copy out of this phase's staging buffer must have completed
  if (B->stage_done[ph]) HIPCHK(hipEventSynchronize(B->stage_done[ph]));
  if (batch_stage_reserve(B, ph, off[ng])) return 1;
  for (size_t q = 0; q < ng; q++) {
    const LaunchRec & r0 = B->rec[groups[q].cases[0]].recs[groups[q].r];
    r0.merge(B, groups[q].r, groups[q].cases, B->h_stage[ph] + off[q], nullptr);
  }
  auto enqueue = [&]() -> int {
    if (pre && pre()) return 1;
    HIPCHK(hipMemcpyAsync(B->d_stage[ph], B->h_stage[ph], off[ng], hipMemcpyHostToDevice, B->stream));
    for (size_t q = 0; q < ng; q++) {
      const LaunchRec & r0 = B->rec[groups[q].cases[0]].recs[groups[q].r];
      if (r0.merge(B, groups[q].r, groups[q].cases, nullptr, B->d_stage[ph] + off[q]) == 0)
        return fail("batch: launch of %s failed", r0.name);
    }
    return post ? post() : 0;
  };
  if (!B->graphs) {
    if (enqueue()) return 1;
  } else {
    std::vector<size_t> sig = {ng, off[ng], (size_t)B->h_stage[ph], (size_t)B->d_stage[ph]};
    for (size_t q = 0; q < ng; q++) {
      const LaunchRec & r0 = B->rec[groups[q].cases[0]].recs[groups[q].r];
      size_t total = 0, shm = 0;
      for (int c : groups[q].cases) {
        const LaunchRec & r = B->rec[c].recs[groups[q].r];
        total += (size_t)r.grid.x * r.grid.y * r.grid.z;
        shm = r.shm > shm ? r.shm : shm;
      }
      const size_t v[] = {(size_t)r0.merge, groups[q].cases.size(), off[q], total, shm, r0.block.x, r0.block.y,
                          r0.block.z};
      sig.insert(sig.end(), v, v + 8);
    }
    if (!B->gexec[ph] || B->gsig[ph] != sig) {
      batch_drop_graph(B, ph);
      hipGraph_t g = nullptr;
      HIPCHK(hipStreamBeginCapture(B->stream, hipStreamCaptureModeThreadLocal));
      const int erc = enqueue();
      const hipError_t ce = hipStreamEndCapture(B->stream, &g);
      if (erc || ce != hipSuccess) {
        if (g) (void)hipGraphDestroy(g);
        return erc ? 1 : fail("batch: graph capture of phase %d failed: %s", ph, hipGetErrorString(ce));
      }
      const hipError_t ie = hipGraphInstantiate(&B->gexec[ph], g, nullptr, nullptr, 0);
      (void)hipGraphDestroy(g);
      if (ie != hipSuccess) {
        B->gexec[ph] = nullptr;
        return fail("batch: graph instantiation of phase %d failed: %s", ph, hipGetErrorString(ie));
      }
      B->gsig[ph] = sig;
      B->graph_captures++;
    }
    HIPCHK(hipGraphLaunch(B->gexec[ph], B->stream));
    B->graph_launches++;
  }
  HIPCHK(hipEventRecord(B->stage_done[ph], B->stream));
  return 0;
}

static int batch_create_one(int ncases, const domain_t * const domains[], fluid_t * const fluids[],
                            interface_t * const interfaces[], vof_batch_t ** out) {
  vof_batch * B = new vof_batch();
  B->graphs = true;
  for (int c = 0; c < ncases; c++) {
    vof_ctx * ctx = ctx_of(domains[c]);
    if (!ctx) {
      delete B;
      return fail("vof_batch_create: case %d has no initialised domain", c);
    }
    if (ctx->schedule != 1 || (ctx->fuse & 63) != 63) {
      delete B;
      return fail("vof_batch_create: case %d is not on the fully fused schedule", c);
    }
    for (int d = 0; d < c; d++)
      if (B->ctx[d] == ctx) {
        delete B;
        return fail("vof_batch_create: case %d repeats case %d", c, d);
      }
    B->ctx.push_back(ctx);
    B->domain.push_back(domains[c]);
    B->fluid.push_back(fluids[c]);
    B->interface.push_back(interfaces[c]);
  }
  B->rec.resize(ncases);
  bool ok = hipStreamCreateWithFlags(&B->stream, hipStreamNonBlocking) == hipSuccess;
  for (int ph = 0; ph < 4 && ok; ph++)
    ok = hipEventCreateWithFlags(&B->stage_done[ph], hipEventDisableTiming) == hipSuccess;
  if (!ok) {
    vof_batch_destroy(B);
    return fail("vof_batch_create: stream / event creation failed");
  }
  int rc = 0;
  // every case's tables in one pinned block and one device block, stage
  // major: the four tables of every case for stage rk are one contiguous run
  size_t tab = 0;
  for (auto * ctx : B->ctx) tab += ctx->tab_stride * 4;
  B->tab_doubles = 3 * tab;
  const size_t tab_stage = tab;
  tab = B->tab_doubles;
  if (hipHostMalloc(&B->h_tab, sizeof(double) * tab) != hipSuccess || hipMalloc(&B->d_tab, sizeof(double) * tab) != hipSuccess)
    rc = fail("vof_batch_create: table allocation failed");
  std::vector<const double *> src(ncases);
  for (int c = 0; c < ncases && !rc; c++) {
    vof_ctx * ctx = B->ctx[c];
    rc = ensure_vof_spare(ctx, B->interface[c]);
    if (!rc && (ctx->fuse & F_PRED)) rc = ensure_field_spares(ctx, B->fluid[c]);
    if (!rc && hipStreamSynchronize(ctx->stream) != hipSuccess) rc = fail("vof_batch_create: case stream failed");
    src[c] = ctx->d_part + ctx->npart;
  }
  if (!rc) {
    size_t o = 0;
    for (auto * ctx : B->ctx) {
      host_free(ctx->h_tab);
      dev_free(ctx->d_tab);
      ctx->h_tab = B->h_tab + o;
      ctx->d_tab = B->d_tab + o;
      ctx->tab_rk = tab_stage;
      ctx->tab_shared = true;
      o += ctx->tab_stride * 4;
    }
    if (hipMalloc(&B->d_dt, sizeof(double) * ncases) != hipSuccess ||
        hipHostMalloc(&B->h_dt, sizeof(double) * ncases) != hipSuccess ||
        hipMalloc(&B->d_dt_src, sizeof(double *) * ncases) != hipSuccess ||
        hipMemcpy(B->d_dt_src, src.data(), sizeof(double *) * ncases, hipMemcpyHostToDevice) != hipSuccess)
      rc = fail("vof_batch_create: dt buffers failed");
  }
  if (rc) {
    vof_batch_destroy(B);
    return 1;
  }
  {
    // up to 8 recording workers (the calling thread records too)
    const unsigned hw = std::thread::hardware_concurrency();
    int nt = (int)(hw > 1 ? hw / 2 : 1);
    nt = nt > 8 ? 8 : nt;
    nt = nt > ncases - 1 ? ncases - 1 : nt;
    if (nt > 0) B->pool.start(nt);
  }
  *out = B;
  return 0;
}

extern "C" int vof_batch_create(int ncases, const domain_t * const domains[], fluid_t * const fluids[],
                                interface_t * const interfaces[], vof_batch_t ** out) {
  if (ncases < 1 || !domains || !fluids || !interfaces || !out) return fail("vof_batch_create: bad arguments");
  const int np = batch_parts(ncases);
  if (np == 1) return batch_create_one(ncases, domains, fluids, interfaces, out);
  vof_batch * B = new vof_batch();
  for (int c = 0; c < ncases; c++) {
    vof_ctx * ctx = ctx_of(domains[c]);
    if (!ctx) {
      delete B;
      return fail("vof_batch_create: case %d has no initialised domain", c);
    }
    for (int d = 0; d < c; d++)
      if (B->ctx[d] == ctx) {
        delete B;
        return fail("vof_batch_create: case %d repeats case %d", c, d);
      }
    B->ctx.push_back(ctx);
    B->domain.push_back(domains[c]);
    B->fluid.push_back(fluids[c]);
    B->interface.push_back(interfaces[c]);
  }
  B->part_cases.resize(np);
  for (int c = 0; c < ncases; c++) B->part_cases[c % np].push_back(c);
  for (int p = 0; p < np; p++) {
    std::vector<const domain_t *> d;
    std::vector<fluid_t *> f;
    std::vector<interface_t *> itf;
    for (int c : B->part_cases[p]) {
      d.push_back(domains[c]);
      f.push_back(fluids[c]);
      itf.push_back(interfaces[c]);
    }
    vof_batch_t * P = nullptr;
    if (batch_create_one((int)d.size(), d.data(), f.data(), itf.data(), &P)) {
      const std::string err = g_err;
      vof_batch_destroy(B);
      g_err = err;
      return 1;
    }
    B->parts.push_back(P);
  }
  *out = B;
  return 0;
}

extern "C" int vof_batch_integrate(vof_batch_t * B, double dts[]) {
  if (!B) return fail("vof_batch_integrate: batch is NULL");
  if (!B->parts.empty()) {
    for (size_t p = 0; p < B->parts.size(); p++) {
      std::vector<double> d(B->part_cases[p].size());
      if (vof_batch_integrate(B->parts[p], d.data())) return 1;
      if (dts)
        for (size_t i = 0; i < d.size(); i++) dts[B->part_cases[p][i]] = d[i];
    }
    return 0;
  }
  const size_t n = B->ctx.size();
  // each case's launches are recorded on a worker thread; an error message
  // (thread-local) is carried back to this thread
  std::vector<int> rcs(n, 0);
  std::vector<std::string> errs(n);
  auto gather_rc = [&]() -> int {
    for (size_t c = 0; c < n; c++)
      if (rcs[c]) {
        g_err = errs[c];
        return 1;
      }
    return 0;
  };
  // dt of every case (decide_dt.c:104-145): batched reductions, one copy back
  B->pool.run((int)n, [&](int c) {
    B->rec[c].recs.clear();
    B->ctx[c]->rec = &B->rec[c];
    rcs[c] = decide_dt_enqueue(B->ctx[c], B->fluid[c]);
    B->ctx[c]->rec = nullptr;
    if (rcs[c]) errs[c] = g_err;
  });
  if (gather_rc()) return 1;
  auto gather = [&]() -> int {
    hipLaunchKernelGGL(k_gather_dt, dim3((unsigned)((n + 63) / 64)), dim3(64), 0, B->stream,
                       (const double * const *)B->d_dt_src, (int)n, B->d_dt);
    LAUNCHCHK();
    HIPCHK(hipMemcpyAsync(B->h_dt, B->d_dt, sizeof(double) * n, hipMemcpyDeviceToHost, B->stream));
    return 0;
  };
  if (batch_run(B, 0, nullptr, gather)) return 1;
  HIPCHK(hipStreamSynchronize(B->stream));
  std::vector<double> dt(n);
  for (size_t c = 0; c < n; c++) {
    B->ctx[c]->h_scalar[0] = B->h_dt[c];
    if (decide_dt_finish(B->ctx[c], B->domain[c], B->fluid[c], &dt[c], false)) return 1;
  }
  // the RK stages of every case, recorded then merged stage by stage: the
  // host records stage rk + 1 while the device runs stage rk
  const size_t tab_stage = B->tab_doubles / 3;
  for (size_t rk = 0; rk < 3; rk++) {
    B->pool.run((int)n, [&](int c) {
      B->rec[c].recs.clear();
      B->ctx[c]->rec = &B->rec[c];
      rcs[c] = integrate_fused_stages(B->domain[c], B->fluid[c], B->interface[c], dt[c], rk, rk + 1);
      B->ctx[c]->rec = nullptr;
      if (rcs[c]) errs[c] = g_err;
    });
    if (gather_rc()) return 1;
    auto tables = [&]() -> int {
      HIPCHK(hipMemcpyAsync(B->d_tab + rk * tab_stage, B->h_tab + rk * tab_stage, sizeof(double) * tab_stage,
                            hipMemcpyHostToDevice, B->stream));
      return 0;
    };
    if (batch_run(B, (int)rk + 1, tables, nullptr)) return 1;
  }
  if (dts)
    for (size_t c = 0; c < n; c++) dts[c] = dt[c];
  return 0;
}

extern "C" int vof_batch_set_graphs(vof_batch_t * B, int on) {
  if (!B) return fail("vof_batch_set_graphs: batch is NULL");
  if (!B->parts.empty()) {
    for (auto * P : B->parts)
      if (vof_batch_set_graphs(P, on)) return 1;
    return 0;
  }
  HIPCHK(hipStreamSynchronize(B->stream));
  B->graphs = on != 0;
  if (!B->graphs)
    for (int ph = 0; ph < 4; ph++) batch_drop_graph(B, ph);
  return 0;
}

extern "C" int vof_batch_graph_counts(const vof_batch_t * B, size_t * captures, size_t * launches) {
  if (!B || !captures || !launches) return fail("vof_batch_graph_counts: NULL argument");
  if (!B->parts.empty()) {
    size_t c = 0, l = 0;
    for (auto * P : B->parts) {
      size_t a = 0, b = 0;
      if (vof_batch_graph_counts(P, &a, &b)) return 1;
      c += a;
      l += b;
    }
    *captures = c;
    *launches = l;
    return 0;
  }
  *captures = B->graph_captures;
  *launches = B->graph_launches;
  return 0;
}

extern "C" int vof_batch_synchronize(vof_batch_t * B) {
  if (!B) return fail("vof_batch_synchronize: batch is NULL");
  if (!B->parts.empty()) {
    for (auto * P : B->parts)
      if (vof_batch_synchronize(P)) return 1;
    return 0;
  }
  HIPCHK(hipStreamSynchronize(B->stream));
  return 0;
}

extern "C" int vof_batch_destroy(vof_batch_t * B) {
  if (!B) return 0;
  if (!B->parts.empty() || !B->part_cases.empty()) {
    int rc = 0;
    for (auto * P : B->parts)
      if (vof_batch_destroy(P)) rc = 1;
    delete B;
    return rc;
  }
  if (B->stream) (void)hipStreamSynchronize(B->stream);
  // cases keep working on their own after the batch: give them private
  // tables.  A case whose tables cannot be allocated is left without any
  // (upload_tab then refuses to step it) and the call reports the failure.
  int rc = 0;
  for (auto * ctx : B->ctx) {
    if (!ctx->tab_shared) continue;
    double * h = nullptr;
    double * d = nullptr;
    if (hipHostMalloc(&h, sizeof(double) * ctx->tab_stride * 12) == hipSuccess &&
        hipMalloc(&d, sizeof(double) * ctx->tab_stride * 12) == hipSuccess) {
      ctx->h_tab = h;
      ctx->d_tab = d;
      ctx->tab_rk = 4 * ctx->tab_stride;
    } else {
      host_free(h);
      ctx->h_tab = nullptr;
      ctx->d_tab = nullptr;
      rc = fail("vof_batch_destroy: no memory for a case's own Thomas tables; that case can no longer be stepped");
    }
    ctx->tab_shared = false;
  }
  for (int ph = 0; ph < 4; ph++) {
    batch_drop_graph(B, ph);
    host_free(B->h_stage[ph]);
    dev_free(B->d_stage[ph]);
    if (B->stage_done[ph]) (void)hipEventDestroy(B->stage_done[ph]);
  }
  host_free(B->h_tab);
  dev_free(B->d_tab);
  dev_free(B->d_dt);
  host_free(B->h_dt);
  dev_free((void *)B->d_dt_src);
  if (B->stream) (void)hipStreamDestroy(B->stream);
  delete B;
  return rc;
}

// ---------------------------------------------------------------------------
// Autotuner.  The tile parameters of vof_set_tuning and the FFT radix are
// picked by timing whole steps of the caller's own cases: coordinate descent
// over the parameters in a fixed order, every candidate timed over `steps`
// steps (after one untimed step that loads its code) from the same state.
// The state of every case is copied on the device first and put back after
// each trial and at the end, so the cases are left as they were.  A
// candidate replaces the current value only when it is at least 1 % faster.

struct TuneSnap {
  std::vector<array_t *> arr;
  std::vector<void *> buf;      // each array's buffer at snapshot time
  std::vector<void *> copy;     // its contents
  std::vector<size_t> bytes;
  vof_ctx * ctx = nullptr;
  double * spare = nullptr;     // the VOF ping-pong buffer and its contents
  void * spare_copy = nullptr;
  array_t * fld[4] = {nullptr, nullptr, nullptr, nullptr};   // ux, uy, uz, t (ping-pong with d_fspare)
  void * fld_buf[4] = {nullptr, nullptr, nullptr, nullptr};  // their buffers at snapshot time
  TuneSnap() = default;
  TuneSnap(const TuneSnap &) = delete;
  TuneSnap & operator=(const TuneSnap &) = delete;
  ~TuneSnap() {
    for (void * c : copy) dev_free(c);
    dev_free(spare_copy);
  }
};

static int tune_take(vof_ctx * ctx, fluid_t * f, interface_t * s, hipStream_t st, TuneSnap & S) {
  std::vector<array_t *> all = {&f->ux, &f->uy, &f->p, &f->psi, &f->t, &s->vof, &s->ifrcx, &s->ifrcy, &s->dvof,
                                &s->normal, &s->curv, &s->flxx, &s->flxy, &s->src[0], &s->src[1]};
#if NDIMS == 3
  all.push_back(&f->uz);
  all.push_back(&s->ifrcz);
  all.push_back(&s->flxz);
  for (int r = 0; r < 3; r++) all.push_back(&f->srcuz[r]);
#endif
  for (int r = 0; r < 3; r++) {
    all.push_back(&f->srcux[r]);
    all.push_back(&f->srcuy[r]);
    all.push_back(&f->srct[r]);
  }
  S.ctx = ctx;
  for (array_t * a : all) {
    if (!a->data || !a->datasize) continue;
    void * c = nullptr;
    HIPCHK(hipMalloc(&c, a->datasize));
    S.copy.push_back(c);
    S.arr.push_back(a);
    S.buf.push_back(a->data);
    S.bytes.push_back(a->datasize);
    HIPCHK(hipMemcpyAsync(c, a->data, a->datasize, hipMemcpyDeviceToDevice, st));
  }
  S.fld[0] = &f->ux;
  S.fld[1] = &f->uy;
#if NDIMS == 3
  S.fld[2] = &f->uz;
#endif
  S.fld[3] = &f->t;
  for (int m = 0; m < 4; m++) S.fld_buf[m] = S.fld[m] ? S.fld[m]->data : nullptr;
  S.spare = ctx->d_vof_spare;
  if (S.spare) {
    HIPCHK(hipMalloc(&S.spare_copy, ctx->vof_bytes));
    HIPCHK(hipMemcpyAsync(S.spare_copy, S.spare, ctx->vof_bytes, hipMemcpyDeviceToDevice, st));
  }
  return 0;
}

static int tune_restore(TuneSnap & S, hipStream_t st) {
  // buffers first (undoing the pointer swaps of the steps), then contents.
  // A field and its second buffer are a pair whose roles the one-kernel
  // predictor swaps: the field gets its snapshot buffer back and the second
  // buffer is the pair's other member (whose contents the predictor rewrites
  // in full before reading them).
  for (int m = 0; m < 4; m++) {
    if (!S.fld[m]) continue;
    void * cur = S.fld[m]->data;
    if (cur != S.fld_buf[m]) S.ctx->d_fspare[m] = (double *)cur;
  }
  for (size_t q = 0; q < S.arr.size(); q++) S.arr[q]->data = S.buf[q];
  S.ctx->d_vof_spare = S.spare;
  for (size_t q = 0; q < S.arr.size(); q++)
    HIPCHK(hipMemcpyAsync(S.buf[q], S.copy[q], S.bytes[q], hipMemcpyDeviceToDevice, st));
  if (S.spare) HIPCHK(hipMemcpyAsync(S.spare, S.spare_copy, S.ctx->vof_bytes, hipMemcpyDeviceToDevice, st));
  return 0;
}

struct TuneKnob {
  const char * name;
  std::vector<int> vals;
  int vof_ctx::*field;
};

// candidate values valid for every case (tile parameters that the cases'
// sizes do not use are left out)
static std::vector<TuneKnob> tune_knobs(const std::vector<vof_ctx *> & cs) {
  bool any256 = false, any_other = false, over1024 = false;
  for (const vof_ctx * c : cs) {
    const int lens[2] = {c->ny, NDIMS == 3 ? c->nz : c->ny};
    for (int L : lens) {
      if (L == 256) any256 = true;
      else any_other = true;
      if (L > 1024) over1024 = true;
    }
  }
  std::vector<TuneKnob> k;
  if (any256) k.push_back({"fft_radix", {16, 4}, &vof_ctx::fft_radix});
  if ((any_other || any256) && !over1024) k.push_back({"fft_elems", {1024, 2048}, &vof_ctx::fft_elems});
#if NDIMS == 3
  k.push_back({"iface_planes", {8, 0, 4, 16}, &vof_ctx::iface_planes});
#endif
  k.push_back({"pred_rows", {8, 4, 2}, &vof_ctx::pred_rows});
  k.push_back({"rhs_rows", {8, 1, 2, 4, 16}, &vof_ctx::rhs_rows});
  k.push_back({"iface_overlap", {1, 2, 0}, &vof_ctx::iface_overlap});
  k.push_back({"thomas_rows", {8, 16, 32}, &vof_ctx::thomas_rows});
  k.push_back({"lines_lpb", {0, 2, 4, 8, 16, 32}, &vof_ctx::lines_lpb});
  return k;
}

// sync_all (may be null): the steps run on several streams (a batch in
// parts); a trial then waits for all of them and is timed on the host clock
// from the end of its untimed step to the end of its last step
static int tune_run(const std::vector<vof_ctx *> & cs, std::vector<std::unique_ptr<TuneSnap>> & snaps,
                    hipStream_t st, int steps, const std::function<int()> & step, char * report, size_t len,
                    const std::function<int()> * sync_all = nullptr) {
  std::string rep;
  char line[160];
  hipEvent_t e0 = nullptr, e1 = nullptr;
  HIPCHK(hipEventCreate(&e0));
  HIPCHK(hipEventCreate(&e1));
  int rc = 0;
  auto trial = [&](double * ms) -> int {
    for (auto & S : snaps)
      if (tune_restore(*S, st)) return 1;
    if (sync_all) {
      if ((*sync_all)()) return 1;
      if (step() || (*sync_all)()) return 1;
      const auto t0 = std::chrono::steady_clock::now();
      for (int n = 0; n < steps; n++)
        if (step()) return 1;
      if ((*sync_all)()) return 1;
      *ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count() / steps;
      return 0;
    }
    if (step()) return 1;
    HIPCHK(hipEventRecord(e0, st));
    for (int n = 0; n < steps; n++)
      if (step()) return 1;
    HIPCHK(hipEventRecord(e1, st));
    HIPCHK(hipEventSynchronize(e1));
    float t = 0.f;
    HIPCHK(hipEventElapsedTime(&t, e0, e1));
    *ms = (double)t / steps;
    return 0;
  };
  std::vector<TuneKnob> knobs = tune_knobs(cs);
  double best = 0.;
  rc = trial(&best);
  if (!rc) {
    snprintf(line, sizeof line, "start - %.4f\n", best);
    rep += line;
  }
  for (const TuneKnob & kn : knobs) {
    if (rc) break;
    const int cur = cs[0]->*kn.field;
    int pick = cur;
    for (int v : kn.vals) {
      if (v == cur) continue;
      for (vof_ctx * c : cs) c->*kn.field = v;
      double ms = 0.;
      if ((rc = trial(&ms))) break;
      snprintf(line, sizeof line, "%s %d %.4f\n", kn.name, v, ms);
      rep += line;
      if (ms < 0.99 * best) {
        best = ms;
        pick = v;
      }
    }
    for (vof_ctx * c : cs) c->*kn.field = pick;
    if (!rc) {
      snprintf(line, sizeof line, "best %s %d %.4f\n", kn.name, pick, best);
      rep += line;
    }
  }
  // leave every case as it was (also after a failed trial: the snapshot is
  // put back whatever the trial left)
  if (rc) {
    if (sync_all)
      (void)(*sync_all)();
    else
      (void)hipStreamSynchronize(st);
  }
  for (auto & S : snaps)
    if (tune_restore(*S, st)) rc = 1;
  if (hipStreamSynchronize(st) != hipSuccess) rc = fail("autotune: stream failed");
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (report && len) {
    strncpy(report, rep.c_str(), len - 1);
    report[len - 1] = 0;
  }
  return rc;
}

extern "C" int vof_autotune(domain_t * domain, fluid_t * fluid, interface_t * interface, int steps, char * report,
                            size_t len) {
  CTX_OR_FAIL(domain);
  if (!fluid || !interface) return fail("vof_autotune: fluid or interface is NULL");
  if (steps < 1 || steps > 1000) return fail("vof_autotune: steps must be in 1..1000");
  if (ctx->schedule != 1 || (ctx->fuse & 63) != 63) return fail("vof_autotune: fused schedule only");
  if (ctx->rec || ctx->tape) return fail("vof_autotune: case is recording");
  if (ensure_vof_spare(ctx, interface)) return 1;
  if ((ctx->fuse & F_PRED) && ensure_field_spares(ctx, fluid)) return 1;
  const bool prof = ctx->prof_on;
  ctx->prof_on = false;
  std::vector<vof_ctx *> cs = {ctx};
  std::vector<std::unique_ptr<TuneSnap>> snaps;
  snaps.emplace_back(new TuneSnap());
  int rc = tune_take(ctx, fluid, interface, ctx->stream, *snaps[0]);
  if (!rc) {
    double dt = 0.;
    rc = tune_run(cs, snaps, ctx->stream, steps, [&]() { return integrate(domain, fluid, interface, &dt); },
                  report, len);
  }
  if (rc) (void)hipStreamSynchronize(ctx->stream);
  ctx->prof_on = prof;
  return rc;
}

extern "C" int vof_batch_autotune(vof_batch_t * B, int steps, char * report, size_t len) {
  if (!B) return fail("vof_batch_autotune: batch is NULL");
  if (steps < 1 || steps > 1000) return fail("vof_batch_autotune: steps must be in 1..1000");
  // a batch in parts is timed as it runs, its parts stepped in turn on their
  // own streams: on the host clock, with every part's stream drained around
  // the timed steps; the snapshots go to the first part's stream
  const bool split = !B->parts.empty();
  hipStream_t st = split ? B->parts[0]->stream : B->stream;
  const std::function<int()> sync_all = [&]() -> int {
    for (auto * P : B->parts) HIPCHK(hipStreamSynchronize(P->stream));
    return 0;
  };
  if (split) {
    if (sync_all()) return 1;
  } else {
    HIPCHK(hipStreamSynchronize(B->stream));
  }
  auto step = [&]() -> int { return vof_batch_integrate(B, nullptr); };
  std::vector<std::unique_ptr<TuneSnap>> snaps;
  std::vector<bool> prof(B->ctx.size());
  int rc = 0;
  for (size_t c = 0; c < B->ctx.size(); c++) {
    prof[c] = B->ctx[c]->prof_on;
    B->ctx[c]->prof_on = false;
  }
  for (size_t c = 0; c < B->ctx.size() && !rc; c++) {
    if (hipStreamSynchronize(B->ctx[c]->stream) != hipSuccess) {
      rc = fail("vof_batch_autotune: stream of case %zu failed", c);
      break;
    }
    snaps.emplace_back(new TuneSnap());
    rc = tune_take(B->ctx[c], B->fluid[c], B->interface[c], st, *snaps.back());
  }
  if (!rc) rc = tune_run(B->ctx, snaps, st, steps, step, report, len, split ? &sync_all : nullptr);
  if (rc) (void)hipStreamSynchronize(st);
  if (split)
    for (auto * P : B->parts) (void)hipStreamSynchronize(P->stream);
  for (size_t c = 0; c < B->ctx.size(); c++) B->ctx[c]->prof_on = prof[c];
  return rc;
}

// decision record / replay of the 2D fused step (the map the adjoint
// differentiates, for its finite-difference checks): mode 1 records the
// decisions of every following RK stage from the first slot on, mode 2
// imposes the recorded ones stage by stage from the first slot on, 0 stops
extern "C" int vof_set_freeze(domain_t * domain, int mode) {
  CTX_OR_FAIL(domain);
#if NDIMS == 2
  if (mode < 0 || mode > 2) return fail("vof_set_freeze: mode %d is not 0, 1 or 2", mode);
  if (mode == 2 && ctx->freeze_n == 0) return fail("vof_set_freeze: no recorded decisions to replay");
  if (mode != 0 && (ctx->schedule != 1 || (ctx->fuse & (F_IFACE | F_VOF)) != (F_IFACE | F_VOF)))
    return fail("vof_set_freeze: needs the fused curvature and VOF kernels");
  if (mode == 1) ctx->freeze_n = 0;
  ctx->freeze_mode = mode;
  ctx->freeze_pos = 0;
  return 0;
#else
  (void)ctx;
  return mode == 0 ? 0 : fail("vof_set_freeze: decision replay is implemented for 2D");
#endif
}

// one step with a given dt (the frozen dt sequence of the adjoint and its
// finite-difference checks); fused schedule
extern "C" int vof_integrate_fixed_dt(const domain_t * domain, fluid_t * fluid, interface_t * interface,
                                      double dt) {
  CTX_OR_FAIL(domain);
  if (ctx->schedule != 1 || (ctx->fuse & 63) != 63) return fail("vof_integrate_fixed_dt: fused schedule only");
  if (ensure_vof_spare(ctx, interface)) return 1;
  if ((ctx->fuse & F_PRED) && ensure_field_spares(ctx, fluid)) return 1;
  return integrate_fused_stages(domain, fluid, interface, dt);
}

#if NDIMS == 2
// ---------------------------------------------------------------------------
// Adjoint of the 2D step (hand-written reverse mode, vof_adjoint.h).
//
// Loss after nsteps: the droplet centroid x = sum xc vof dV / sum vof dV
// plus the interface length sum |grad vof| dV (central differences, x walls
// at vof 0, periodic y; |grad vof| regularised below vof differences of 1e-6
// across a cell).  The map differentiated is the step with its discrete
// decisions held (upwind sides, single-phase tests, Newton update counts:
// vof_set_freeze replays them) and with the corner normals of flat corners
// (|grad vof| DXC below the floor, 1e-2 by default) held constant.  The
// forward runs with the given dt sequence and
// keeps the state (ux, uy, p, vof) every `every` steps; each segment is then
// recomputed with its stage states recorded and reversed stage by stage.
// The Poisson and x-line solves are transposed as D S D^-1 around the forward
// solvers S (the operators are D^-1 K with K symmetric, D the cell widths).

static constexpr double kLenFloor = 1e-6;

static double loss_of(const vof_ctx * ctx, const domain_t * domain, const std::vector<double> & vof,
                      std::vector<double> * grad) {
  const int nx = ctx->nx, ny = ctx->ny;
  const double dy = ctx->dy;
  const double * xc = domain->xc;
  const double * dxf = domain->dxf;
  const double * dxc = domain->dxc;
  auto V = [&](int i, int j) -> double {
    if (i < 1 || i > nx) return 0.;
    j = j < 1 ? j + ny : (j > ny ? j - ny : j);
    return vof[I_VOF(i, j, 1)];
  };
  // compensated sums: the loss of a perturbed run is compared with the
  // unperturbed one to ~1e-16 relative by finite differences
  struct Sum {
    double s = 0., c = 0.;
    void add(double v) {
      const double y = v - c, t = s + y;
      c = (t - s) - y;
      s = t;
    }
  };
  Sum s0, s1, len;
  for (int j = 1; j <= ny; j++)
    for (int i = 1; i <= nx; i++) {
      const double dv = dxf[i - 1] * dy;
      s0.add(xc[i] * V(i, j) * dv);
      s1.add(V(i, j) * dv);
    }
  const double X = s0.s / s1.s;
  if (grad) grad->assign(vof.size(), 0.);
  auto G = [&](int i, int j, double v) {
    if (i < 1 || i > nx) return;
    j = j < 1 ? j + ny : (j > ny ? j - ny : j);
    (*grad)[I_VOF(i, j, 1)] += v;
  };
  for (int j = 1; j <= ny; j++)
    for (int i = 1; i <= nx; i++) {
      const double dv = dxf[i - 1] * dy;
      const double hx = dxc[i - 1] + dxc[i];
      const double gx = (V(i + 1, j) - V(i - 1, j)) / hx;
      const double gy = (V(i, j + 1) - V(i, j - 1)) / (2. * dy);
      // |grad vof| regularised below g0 (vof differences of 1e-6 across the
      // cell): sqrt(|g|^2 + g0^2) - g0 is |g| - g0 + O(g0^2 / |g|) at the
      // interface and |g|^2 / (2 g0) in single-phase regions, whose
      // rounding-level vof variations then add nothing measurable (and no
      // kink) to the length
      const double g0 = kLenFloor / hx;
      const double nrm = sqrt(gx * gx + gy * gy + g0 * g0);
      len.add((nrm - g0) * dv);
      if (grad) {
        G(i, j, dv * (xc[i] - X) / s1.s);
        G(i + 1, j, dv * gx / nrm / hx);
        G(i - 1, j, -dv * gx / nrm / hx);
        G(i, j + 1, dv * gy / nrm / (2. * dy));
        G(i, j - 1, -dv * gy / nrm / (2. * dy));
      }
    }
  return X + len.s;
}

extern "C" int vof_adjoint_loss(const domain_t * domain, const interface_t * interface, double * loss) {
  CTX_OR_FAIL(domain);
  std::vector<double> h(interface->vof.datasize / sizeof(double));
  HIPCHK(hipMemcpyAsync(h.data(), interface->vof.data, interface->vof.datasize, hipMemcpyDeviceToHost,
                        ctx->stream));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  *loss = loss_of(ctx, domain, h, nullptr);
  return 0;
}

struct AdjBuf {
  std::vector<void *> all;
  int get(double ** p, size_t bytes) {
    if (hipMalloc(p, bytes) != hipSuccess) return fail("adjoint: allocation of %zu bytes failed", bytes);
    all.push_back(*p);
    return 0;
  }
  ~AdjBuf() {
    for (void * p : all) dev_free(p);
  }
};

extern "C" int vof_adjoint_gradient(const domain_t * domain, fluid_t * fluid, interface_t * interface, int nsteps,
                                    const double * dts, int every, double * loss, double * grad_ux,
                                    double * grad_uy) {
  CTX_OR_FAIL(domain);
  if (nsteps < 1 || !dts || every < 1) return fail("vof_adjoint_gradient: bad arguments");
  if (ctx->schedule != 1 || (ctx->fuse & 63) != 63) return fail("vof_adjoint_gradient: fused schedule only");
  {
    double buoy = 0.;
    if (param_get(ctx, P_BUOY, &buoy)) return 1;
    if (buoy != 0.) return fail("vof_adjoint_gradient: buoyancy (temperature coupling) is not differentiated");
  }
  if (ctx->poisson_refine >= 0) return fail("vof_adjoint_gradient: plain FP64 Poisson only");
  if (ctx->freeze_mode) return fail("vof_adjoint_gradient: decision record / replay is on (vof_set_freeze 0 first)");
  if (ensure_vof_spare(ctx, interface)) return 1;
  if ((ctx->fuse & F_PRED) && ensure_field_spares(ctx, fluid)) return 1;
  const int nx = ctx->nx, ny = ctx->ny;
  const size_t bvof = interface->vof.datasize, bux = fluid->ux.datasize, buy = fluid->uy.datasize;
  const size_t bp = fluid->p.datasize;
  const size_t ncell = (size_t)nx * ny;
  hipStream_t st = ctx->stream;
  AdjBuf B;
  // checkpoints
  const int nck = (nsteps + every - 1) / every;
  std::vector<double *> ck(4 * (size_t)nck);
  for (int c = 0; c < nck; c++) {
    if (B.get(&ck[4 * c + 0], bvof) || B.get(&ck[4 * c + 1], bux) || B.get(&ck[4 * c + 2], buy) ||
        B.get(&ck[4 * c + 3], bp))
      return 1;
  }
  auto save = [&](int c) -> int {
    HIPCHK(hipMemcpyAsync(ck[4 * c + 0], interface->vof.data, bvof, hipMemcpyDeviceToDevice, st));
    HIPCHK(hipMemcpyAsync(ck[4 * c + 1], fluid->ux.data, bux, hipMemcpyDeviceToDevice, st));
    HIPCHK(hipMemcpyAsync(ck[4 * c + 2], fluid->uy.data, buy, hipMemcpyDeviceToDevice, st));
    HIPCHK(hipMemcpyAsync(ck[4 * c + 3], fluid->p.data, bp, hipMemcpyDeviceToDevice, st));
    return 0;
  };
  auto restore = [&](int c) -> int {
    HIPCHK(hipMemcpyAsync(interface->vof.data, ck[4 * c + 0], bvof, hipMemcpyDeviceToDevice, st));
    HIPCHK(hipMemcpyAsync(fluid->ux.data, ck[4 * c + 1], bux, hipMemcpyDeviceToDevice, st));
    HIPCHK(hipMemcpyAsync(fluid->uy.data, ck[4 * c + 2], buy, hipMemcpyDeviceToDevice, st));
    HIPCHK(hipMemcpyAsync(fluid->p.data, ck[4 * c + 3], bp, hipMemcpyDeviceToDevice, st));
    return 0;
  };
  // forward with checkpoints
  for (int n = 0; n < nsteps; n++) {
    if (n % every == 0 && save(n / every)) return 1;
    if (integrate_fused_stages(domain, fluid, interface, dts[n])) return 1;
  }
  // loss and its vof gradient on the host
  std::vector<double> hv(bvof / sizeof(double)), hg;
  HIPCHK(hipMemcpyAsync(hv.data(), interface->vof.data, bvof, hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  const double L = loss_of(ctx, domain, hv, &hg);
  // adjoint state and scratch
  double *lux, *luy, *lp, *lvof, *lpsi, *lcurv, *lm, *lnrm, *lfx, *lfy, *ls, *csx, *csy, *cs, *qx, *qy, *vx, *vy,
      *b, *x;
  const size_t bcurv = interface->curv.datasize, bdvof = interface->dvof.datasize;
  const size_t bnrm = interface->normal.datasize;
  if (B.get(&lux, bux) || B.get(&luy, buy) || B.get(&lp, bp) || B.get(&lvof, bvof) || B.get(&lpsi, bp) ||
      B.get(&lcurv, bcurv) || B.get(&lm, bdvof) || B.get(&lnrm, bnrm) ||
      B.get(&lfx, sizeof(double) * (size_t)(nx - 1) * ny) || B.get(&lfy, sizeof(double) * ncell) ||
      B.get(&ls, sizeof(double) * ncell) || B.get(&csx, sizeof(double) * (size_t)(nx - 1) * ny) ||
      B.get(&csy, sizeof(double) * ncell) || B.get(&cs, sizeof(double) * ncell) || B.get(&qx, bux) ||
      B.get(&qy, buy) || B.get(&vx, sizeof(double) * (size_t)(nx - 1) * ny) || B.get(&vy, sizeof(double) * ncell) ||
      B.get(&b, sizeof(double) * ncell) || B.get(&x, bp))
    return 1;
  HIPCHK(hipMemsetAsync(lux, 0, bux, st));
  HIPCHK(hipMemsetAsync(luy, 0, buy, st));
  HIPCHK(hipMemsetAsync(lp, 0, bp, st));
  HIPCHK(hipMemsetAsync(csx, 0, sizeof(double) * (size_t)(nx - 1) * ny, st));
  HIPCHK(hipMemsetAsync(csy, 0, sizeof(double) * ncell, st));
  HIPCHK(hipMemsetAsync(cs, 0, sizeof(double) * ncell, st));
  HIPCHK(hipMemcpyAsync(lvof, hg.data(), bvof, hipMemcpyHostToDevice, st));
  // tapes of one segment: (vof, ux, uy) per stage
  std::vector<double *> tape(9 * (size_t)every);
  for (size_t t = 0; t < tape.size(); t += 3) {
    if (B.get(&tape[t + 0], bvof) || B.get(&tape[t + 1], bux) || B.get(&tape[t + 2], buy)) return 1;
  }
  const dim3 blk(64, 4);
  const dim3 gc((unsigned)((nx + 2 + 63) / 64), (unsigned)((ny + 3) / 4));
  const Grid g = ctx->grid;
  const double m_dif = fluid->m_dif;
  const int NC = 1024 / ny > 0 ? 1024 / ny : 1;
  const int nbx = (nx + 2 * NC - 1) / (2 * NC);
  const int use1024 = ny <= 1024;
  for (int c = nck - 1; c >= 0; c--) {
    const int s0 = c * every, s1 = s0 + every < nsteps ? s0 + every : nsteps;
    if (restore(c)) return 1;
    for (int n = s0; n < s1; n++) {
      ctx->tape = &tape[9 * (size_t)(n - s0)];
      const int rc = integrate_fused_stages(domain, fluid, interface, dts[n]);
      ctx->tape = nullptr;
      if (rc) return 1;
    }
    for (int n = s1 - 1; n >= s0; n--) {
      const double dt = dts[n];
      for (int rk = 2; rk >= 0; rk--) {
        double ** t = &tape[9 * (size_t)(n - s0) + 3 * rk];
        const double * vof = t[0];
        const double * ux = t[1];
        const double * uy = t[2];
        const double ca = kRk[rk][0] * dt, cb = kRk[rk][1] * dt, cg = kRk[rk][2] * dt;
        const double pc = kRk[rk][2] * dt;
        const double pp = 0.5 * kRk[rk][2] * dt * m_dif;
        const double pref = 1. / (kRk[rk][2] * dt);
        // forward quantities of the stage: corner normals, curvature, normals
        hipLaunchKernelGGL(vofk::k_gradient, grid3(nx + 1, ny + 3, 1, kBlk), kBlk, 0, st, g, vof,
                           (double *)interface->dvof.data);
        hipLaunchKernelGGL(vofk::k_normal_curv, grid3(nx, ny + 2, 1, kBlk), kBlk, 0, st, g, vof,
                           (const double *)interface->dvof.data, (double *)interface->normal.data,
                           (double *)interface->curv.data);
        LAUNCHCHK();
        // R1 correction / pressure
        HIPCHK(hipMemsetAsync(lpsi, 0, bp, st));
        hipLaunchKernelGGL(vofa::ka_corr, gc, blk, 0, st, g, pc, pp, (const double *)lux, (const double *)luy,
                           (const double *)lp, lpsi);
        // R2 Poisson: x = G D^-1 lambda_psi, lambda_u* += pref div^T D x
        hipLaunchKernelGGL(vofa::ka_psi_rhs, gc, blk, 0, st, g, (const double *)lpsi, b);
        if (use1024) {
          hipLaunchKernelGGL(vofa::ka_vec_rfft_y<1024>, dim3(nbx), dim3(vofk::kFftThreads), 0, st, g, ctx->plan_y,
                             (const double2 *)ctx->d_twy, (const double *)b, 1. / (double)ny, ctx->d_spec);
        } else {
          hipLaunchKernelGGL(vofa::ka_vec_rfft_y<2048>, dim3((nx + 2 * (2048 / ny) - 1) / (2 * (2048 / ny))),
                             dim3(vofk::kFftThreads), 0, st, g, ctx->plan_y, (const double2 *)ctx->d_twy,
                             (const double *)b, 1. / (double)ny, ctx->d_spec);
        }
        if (ctx->thomas_part && nx >= 4 && nx <= 256) {
          hipLaunchKernelGGL(vofk::k_thomas_part<double2>, dim3((ctx->nh + vofk::kTpLines - 1) / vofk::kTpLines),
                             dim3(vofk::kTpLines * vofk::kTpSeg), 0, st, nx, ctx->nh, ctx->nh,
                             (const double *)ctx->d_pl, (const double *)ctx->d_pu, (const double *)ctx->d_ey,
                             (const double *)ctx->d_ez, (const double *)ctx->d_stab, ctx->d_spec, ctx->d_vbuf);
        } else {
          hipLaunchKernelGGL(vofk::k_thomas_spec<double2>, dim3((2 * ctx->nh + 63) / 64), dim3(64), 0, st, nx, ctx->nh,
                             ctx->nh, (const double *)ctx->d_pl, (const double *)ctx->d_pu,
                             (const double *)ctx->d_ey, (const double *)ctx->d_ez, ctx->d_spec, ctx->d_vbuf,
                             vbuf_store(ctx));
        }
        if (use1024) {
          hipLaunchKernelGGL(vofk::k_irfft_y_psi<1024>, dim3(nbx), dim3(vofk::kFftThreads), 0, st, g, ctx->plan_y,
                             (const double2 *)ctx->d_twy, (const double2 *)ctx->d_spec, x, 1);
        } else {
          hipLaunchKernelGGL(vofk::k_irfft_y_psi<2048>, dim3((nx + 2 * (2048 / ny) - 1) / (2 * (2048 / ny))),
                             dim3(vofk::kFftThreads), 0, st, g, ctx->plan_y, (const double2 *)ctx->d_twy,
                             (const double2 *)ctx->d_spec, x, 1);
        }
        hipLaunchKernelGGL(vofa::ka_div_t, gc, blk, 0, st, g, pref, (const double *)x, lux, luy);
        LAUNCHCHK();
        // R3 x-line solves: q = M^-1 D^-1 lambda_u*
        hipLaunchKernelGGL(vofa::ka_lines_in, gc, blk, 0, st, g, (const double *)lux, (const double *)luy, vx, vy);
        HIPCHK(hipMemsetAsync(qx, 0, bux, st));
        HIPCHK(hipMemsetAsync(qy, 0, buy, st));
        {
          vofk::LineJob lj{};
          lj.nlines = ny;
          lj.ny = ny;
          lj.n1 = 0;
          lj.images = 0;
          lj.nz = 1;
          const double pm = 0.5 * kRk[rk][2] * dt * m_dif;
          const double * tx = nullptr;
          const double * ty = nullptr;
          if (upload_tab(ctx, 0, rk, pm, &tx) || upload_tab(ctx, 1, rk, pm, &ty)) return 1;
          lj.f[0] = vofk::LineField{tx, nullptr, nullptr, nullptr, vx, qx, nx - 1, nx + 1, 1};
          lj.f[1] = vofk::LineField{ty, nullptr, nullptr, nullptr, vy, qy, nx, nx + 2, 1};
          int lpb = (int)(32768 / (sizeof(double) * (nx | 1)));
          lpb = lpb > 64 ? 64 : (lpb < 1 ? 1 : lpb);
          const size_t shm = sizeof(double) * (size_t)lpb * (nx | 1);
          hipLaunchKernelGGL(vofk::k_lines_lds, dim3((unsigned)((ny + lpb - 1) / lpb), 2), dim3(vofk::kWThreads), shm,
                             st, lj, lpb);
          LAUNCHCHK();
        }
        // R7 vof source (before anything else adds to lambda_vof)
        hipLaunchKernelGGL(vofa::ka_vof_src, gc, blk, 0, st, g, kRk[rk][0] * dt, kRk[rk][1] * dt,
                           (const double *)lvof, ls, cs);
        // R4 + R5 predictor sources
        {
          vofa::AdjRhs a{};
          a.ux = ux;
          a.uy = uy;
          a.lapx_ux = ctx->d_lapx_ux;
          a.lapx_c = ctx->d_lapx_c;
          a.qx = qx;
          a.qy = qy;
          a.lux = lux;
          a.luy = luy;
          a.lp = lp;
          a.csx = csx;
          a.csy = csy;
          a.lfx = lfx;
          a.lfy = lfy;
          a.ca = ca;
          a.cb = cb;
          a.cg = cg;
          a.dif = m_dif;
          hipLaunchKernelGGL(vofa::ka_rhs, gc, blk, 0, st, g, a);
        }
        // R8 fluxes, R6 force, R9-R11 interface geometry
        HIPCHK(hipMemsetAsync(lnrm, 0, bnrm, st));
        HIPCHK(hipMemsetAsync(lcurv, 0, bcurv, st));
        HIPCHK(hipMemsetAsync(lm, 0, bdvof, st));
        hipLaunchKernelGGL(vofa::ka_flux, gc, blk, 0, st, g, ux, uy, vof, (const double *)interface->normal.data,
                           (const double *)ls, lux, luy, lvof, lnrm);
        hipLaunchKernelGGL(vofa::ka_force, gc, blk, 0, st, g, interface->tension, vof,
                           (const double *)interface->curv.data, (const double *)lfx, (const double *)lfy, lvof,
                           lcurv);
        hipLaunchKernelGGL(vofa::ka_normal_curv, gc, blk, 0, st, g, vof, (const double *)interface->dvof.data,
                           (const double *)interface->normal.data, (const double *)lnrm, (const double *)lcurv,
                           lvof, lm);
        hipLaunchKernelGGL(vofa::ka_gradient, dim3((unsigned)((nx + 1 + 63) / 64), (unsigned)((ny + 3) / 4)), blk,
                           0, st, g, ctx->adj_floor, vof, (const double *)lm, lvof);
        LAUNCHCHK();
      }
    }
  }
  // the solver is left at its initial state; the gradient is lambda_u
  if (restore(0)) return 1;
  HIPCHK(hipMemcpyAsync(grad_ux, lux, bux, hipMemcpyDeviceToHost, st));
  HIPCHK(hipMemcpyAsync(grad_uy, luy, buy, hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  *loss = L;
  return 0;
}
#endif

// stages of the fused schedule (bit mask; the default is all of them)
extern "C" int vof_set_fusion(domain_t * domain, int mask) {
  CTX_OR_FAIL(domain);
  if (mask < 0 || mask > 127) return fail("vof_set_fusion: mask %d out of range", mask);
  ctx->fuse = mask;
  return 0;
}

// ---------------------------------------------------------------------------
// diagnostics

extern "C" int vof_diagnostics(const domain_t * domain, const fluid_t * fluid,
                               const interface_t * interface, double * out) {
  CTX_OR_FAIL(domain);
  const int nb = 256;
#if NDIMS == 3
  const double * uz = (const double *)fluid->uz.data;
#else
  const double * uz = nullptr;
#endif
  KLN(ctx, "k_diag_partial", vofk::k_diag_partial, dim3(nb), dim3(vofk::kRedThreads),
                     ctx->grid, (const double *)fluid->ux.data, (const double *)fluid->uy.data, uz,
                     (const double *)fluid->t.data, (const double *)interface->vof.data, ctx->d_part);
  LAUNCHCHK();
  std::vector<double> h((size_t)nb * vofk::kDiagN);
  HIPCHK(hipMemcpyAsync(h.data(), ctx->d_part, sizeof(double) * h.size(), hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  double q[vofk::kDiagN];
  for (int m = 0; m < vofk::kDiagN; m++) q[m] = m == 1 ? 1. : 0.;
  for (int b = 0; b < nb; b++) {
    const double * r = &h[(size_t)b * vofk::kDiagN];
    q[0] = fmax(q[0], r[0]);
    q[1] = fmin(q[1], r[1]);
    q[2] = fmax(q[2], r[2]);
    for (int m = 3; m < vofk::kDiagN; m++) q[m] += r[m];
  }
  int o = 0;
  out[o++] = q[0];
  out[o++] = q[1];
  out[o++] = q[2];
  out[o++] = q[3] / q[4];
  out[o++] = q[5];
  out[o++] = q[6];
#if NDIMS == 3
  out[o++] = q[7];
#endif
  out[o++] = q[8];
  out[o++] = q[9];
  out[o++] = q[10];
#if NDIMS == 3
  out[o++] = q[11];
#endif
  out[o++] = q[12];
  return 0;
}

// ---------------------------------------------------------------------------
// Nusselt numbers (logging/nusselt/main.c:18-50): {wall heat flux, kinetic
// energy injection, kinetic energy dissipation, thermal dissipation}, each
// normalised by the conductive flux t_dif * ly (* lz)

extern "C" int vof_nusselt(const domain_t * domain, const fluid_t * fluid, double out[4]) {
  CTX_OR_FAIL(domain);
  const int nb = 256;
#if NDIMS == 3
  const double * uz = (const double *)fluid->uz.data;
  const double ref = fluid->t_dif * domain->lengths[1] * domain->lengths[2];
#else
  const double * uz = nullptr;
  const double ref = fluid->t_dif * domain->lengths[1];
#endif
  KLN(ctx, "k_nusselt_partial", vofk::k_nusselt_partial, dim3(nb), dim3(vofk::kRedThreads), ctx->grid,
      (const double *)fluid->ux.data, (const double *)fluid->uy.data, uz, (const double *)fluid->t.data,
      fluid->m_dif, fluid->t_dif, ctx->d_part);
  std::vector<double> h((size_t)nb * vofk::kNuN);
  HIPCHK(hipMemcpyAsync(h.data(), ctx->d_part, sizeof(double) * h.size(), hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  double q[vofk::kNuN] = {0., 0., 0., 0.};
  for (int b = 0; b < nb; b++)
    for (int m = 0; m < vofk::kNuN; m++) q[m] += h[(size_t)b * vofk::kNuN + m];
  out[0] = q[0] / ref;
  out[1] = q[1] / ref + 1.;
  out[2] = q[2] / ref + 1.;
  out[3] = q[3] / ref;
  return 0;
}

// ---------------------------------------------------------------------------
// statistics (statistics.c): device accumulators of ux, ux^2, uy, uy^2, t,
// t^2, ux t (and uz, uz^2 in 3D), reduced over y (and z) on output.  Like the
// reference module this is one collector per process: get_next_time reports
// the domain that ran init last.

static const char * const kStatNames[9] = {"ux1", "ux2", "uy1", "uy2", "t1", "t2", "uxt", "uz1", "uz2"};
static double g_stat_next = 0.;

static int stat_width(const vof_ctx * ctx, int m) {
  return (m == 0 || m == 1 || m == 6) ? ctx->nx + 1 : ctx->nx + 2;
}

static int stats_init(const domain_t * domain, const double time) {
  CTX_OR_FAIL(domain);
  // config.get_double("stat_rate" / "stat_after") (statistics.c:63-70)
  double rate = 0., after = 0.;
  for (int m = 0; m < 2; m++) {
    const char * name = m == 0 ? "stat_rate" : "stat_after";
    const char * sv = getenv(name);
    if (!sv) return fail("%s not found", name);
    errno = 0;
    const double v = strtod(sv, nullptr);
    if (errno != 0) return fail("%s: invalid value as double", name);
    (m == 0 ? rate : after) = v;
  }
  if (!(rate > 0.)) return fail("stat_rate must be positive");
  ctx->stat_rate = rate;
  ctx->stat_next = rate * ceil(fmax(DBL_EPSILON, fmax(time, after)) / rate);
  g_stat_next = ctx->stat_next;
  const size_t rows = (size_t)ctx->ny * (size_t)ctx->nz;
  for (int m = 0; m < 9; m++) {
    if (NDIMS == 2 && m >= 7) continue;
    const size_t bytes = sizeof(double) * (size_t)stat_width(ctx, m) * rows;
    if (!ctx->d_stat[m]) HIPCHK(hipMalloc(&ctx->d_stat[m], bytes));
    HIPCHK(hipMemsetAsync(ctx->d_stat[m], 0, bytes, ctx->stream));
  }
  ctx->stat_num = 0;
  return 0;
}

static int stats_collect(const domain_t * domain, const fluid_t * fluid) {
  CTX_OR_FAIL(domain);
  if (!ctx->d_stat[0]) return fail("statistics.collect before statistics.init");
  vofk::StatAcc acc;
  for (int m = 0; m < 9; m++) acc.a[m] = ctx->d_stat[m];
#if NDIMS == 3
  const double * uz = (const double *)fluid->uz.data;
#else
  const double * uz = nullptr;
#endif
  KLN(ctx, "k_stats_collect", vofk::k_stats_collect, grid3(ctx->nx + 2, ctx->ny, ctx->nz, kBlk), kBlk,
      ctx->grid, (const double *)fluid->ux.data, (const double *)fluid->uy.data, uz,
      (const double *)fluid->t.data, acc);
  ctx->stat_m_dif = fluid->m_dif;
  ctx->stat_t_dif = fluid->t_dif;
  ctx->stat_num += 1;
  ctx->stat_next += ctx->stat_rate;
  g_stat_next = ctx->stat_next;
  return 0;
}

static int mkdir_p(const char * dir) {
  std::string p(dir);
  for (size_t pos = 1; pos <= p.size(); pos++) {
    if (pos == p.size() || p[pos] == '/') {
      const std::string sub = p.substr(0, pos);
      if (mkdir(sub.c_str(), 0777) != 0 && errno != EEXIST) return fail("mkdir %s failed", sub.c_str());
    }
  }
  return 0;
}

// write the collected statistics into dirname (statistics.c:285-330)
extern "C" int vof_statistics_output_to(const domain_t * domain, const char dirname[]) {
  CTX_OR_FAIL(domain);
  if (ctx->stat_num == 0) return 0;
  if (mkdir_p(dirname)) return 1;
  if (npy::write(dirname, "num", kNpySizeT, {}, &ctx->stat_num, sizeof(size_t))) return 1;
  if (npy::write(dirname, "m_dif", kNpyDouble, {}, &ctx->stat_m_dif, sizeof(double))) return 1;
  if (npy::write(dirname, "t_dif", kNpyDouble, {}, &ctx->stat_t_dif, sizeof(double))) return 1;
  if (domain_save(dirname, domain)) return 1;
  const size_t rows = (size_t)ctx->ny * (size_t)ctx->nz;
  double * d_vec = ctx->d_part;   // nx + 2 <= npart * kDiagN entries
  if ((size_t)(ctx->nx + 2) > (size_t)ctx->npart * vofk::kDiagN) return fail("statistics: profile too long");
  std::vector<double> h((size_t)ctx->nx + 2);
  for (int m = 0; m < 9; m++) {
    if (NDIMS == 2 && m >= 7) continue;
    const int w = stat_width(ctx, m);
    KLN(ctx, "k_stats_reduce", vofk::k_stats_reduce, dim3((w + 63) / 64), dim3(64), ctx->d_stat[m], w, rows, d_vec);
    HIPCHK(hipMemcpyAsync(h.data(), d_vec, sizeof(double) * (size_t)w, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    if (npy::write(dirname, kStatNames[m], kNpyDouble, {(size_t)w}, h.data(), sizeof(double))) return 1;
  }
  return 0;
}

static int stats_output(const domain_t * domain, const size_t step) {
  // statistics.c:291-299: output/stat/step%010zu
  char dir[64];
  snprintf(dir, sizeof dir, "output/stat/step%010zu", step);
  return vof_statistics_output_to(domain, dir);
}

static double stats_get_next_time(void) { return g_stat_next; }

// host-only object: the device pass must not see a table of host functions
#if !defined(__HIP_DEVICE_COMPILE__)
extern "C" const statistics_t statistics = {
    stats_init,
    stats_collect,
    stats_output,
    stats_get_next_time,
};
#endif

// ---------------------------------------------------------------------------
// step / time scalars of an IC or save directory (src/main.c:64-71, 26-27)

extern "C" int vof_read_step_time(const char dirname[], size_t * step, double * time) {
  if (npy::read(dirname, "step", kNpySizeT, {}, step, sizeof(size_t))) return 1;
  if (npy::read(dirname, "time", kNpyDouble, {}, time, sizeof(double))) return 1;
  return 0;
}

extern "C" int vof_write_step_time(const char dirname[], size_t step, double time) {
  if (npy::write(dirname, "step", kNpySizeT, {}, &step, sizeof(size_t))) return 1;
  if (npy::write(dirname, "time", kNpyDouble, {}, &time, sizeof(double))) return 1;
  return 0;
}

// ---------------------------------------------------------------------------
// per-kernel timing and the copy probe

extern "C" int vof_profile(domain_t * domain, int on) {
  CTX_OR_FAIL(domain);
  if (on) {
    if (ctx->ev_pool.empty()) {
      ctx->ev_pool.resize(8192);
      for (auto & e : ctx->ev_pool) HIPCHK(hipEventCreate(&e));
      ctx->ev_next = 0;
    }
    ctx->prof_on = true;
  } else {
    if (prof_flush(ctx)) return 1;
    ctx->prof_on = false;
  }
  return 0;
}

extern "C" int vof_profile_reset(domain_t * domain) {
  CTX_OR_FAIL(domain);
  if (prof_flush(ctx)) return 1;
  ctx->prof_names.clear();
  ctx->prof_ms.clear();
  ctx->prof_cnt.clear();
  return 0;
}

// "name launches total_ms" lines
extern "C" int vof_profile_report(domain_t * domain, char * buf, size_t len) {
  CTX_OR_FAIL(domain);
  if (prof_flush(ctx)) return 1;
  std::string out;
  for (size_t m = 0; m < ctx->prof_names.size(); m++) {
    char line[256];
    snprintf(line, sizeof line, "%s %ld %.9f\n", ctx->prof_names[m].c_str(), ctx->prof_cnt[m], ctx->prof_ms[m]);
    out += line;
  }
  if (out.size() + 1 > len) return fail("vof_profile_report: buffer too small (%zu bytes needed)", out.size() + 1);
  memcpy(buf, out.c_str(), out.size() + 1);
  return 0;
}

// device copy of n doubles, `reps` times; returns the mean (read + write)
// bandwidth in GB/s.  Calibrates hardware byte counters against a kernel
// of known traffic.
extern "C" int vof_copy_probe(domain_t * domain, size_t n, int reps, double * gbps) {
  CTX_OR_FAIL(domain);
  double * a = nullptr;
  double * b = nullptr;
  HIPCHK(hipMalloc(&a, sizeof(double) * n));
  HIPCHK(hipMalloc(&b, sizeof(double) * n));
  HIPCHK(hipMemsetAsync(a, 0, sizeof(double) * n, ctx->stream));
  hipEvent_t e0, e1;
  HIPCHK(hipEventCreate(&e0));
  HIPCHK(hipEventCreate(&e1));
  const unsigned nb = 4096;
  KLN(ctx, "k_copy", vofk::k_copy, dim3(nb), dim3(256), (const double *)a, b, n);
  HIPCHK(hipEventRecord(e0, ctx->stream));
  for (int r = 0; r < reps; r++) KLN(ctx, "k_copy", vofk::k_copy, dim3(nb), dim3(256), (const double *)a, b, n);
  HIPCHK(hipEventRecord(e1, ctx->stream));
  HIPCHK(hipEventSynchronize(e1));
  float ms = 0.f;
  HIPCHK(hipEventElapsedTime(&ms, e0, e1));
  *gbps = 2. * sizeof(double) * (double)n * reps / (ms * 1e-3) / 1e9;
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  dev_free(a);
  dev_free(b);
  return 0;
}

// ---------------------------------------------------------------------------
// tri-diagonal solver vtable (tdm.c:110-324) over device right-hand sides
//
// The reference's gtsv (tdm.c:17-50) recomputes, for every right-hand side,
// the same factors v[i] = u[i] / (c[i] - l[i] v[i-1]) and the reciprocals
// 1 / (c[i] - l[i] v[i-1]); for a periodic system it solves the correction
// vector q1 and the denominator of the Sherman-Morrison update once per call
// (tdm.c:78-86,93).  Here k_tdm_table computes all of these once per call on
// the device, with the reference's operations in the reference's order, and
// k_tdm_lines runs only the per-line recurrences, so every line gets the
// reference's values bit for bit.

struct tdm_info_t_ {
  int size, nrhs;
  bool is_periodic, is_complex;
  double * l, * c, * u;        // host, filled by the caller through get_l / get_c / get_u
  double * d_luc;              // device copy of l, c, u
  double * d_tab;              // factor table (tdm_table_len(size) doubles)
  hipStream_t stream;          // default priority, blocking w.r.t. the null stream
};

namespace vofk {
// factor table of one system of `n` rows:
//   [0, n)        v[i]   (upper factors, rows 0 .. m-2 of the gtsv part)
//   [n, 2n)       rec[i] (reciprocal pivots of rows 1 .. m-1; rec[0] = c[0], a divisor)
//   [2n, 3n)      q1[i]  (periodic: the solved correction vector, rows 0 .. n-2)
//   3n            1 if the last gtsv pivot is singular (|pivot| <= DBL_EPSILON, tdm.c:39-44)
//   3n + 1        den of the Sherman-Morrison update (tdm.c:93)
//   3n + 2        1 if |den| < DBL_EPSILON (tdm.c:94)
// m = n for a non-periodic system, n - 1 for a periodic one
__host__ __device__ constexpr size_t tdm_table_len(int n) { return 3 * (size_t)n + 3; }

// forward and backward sweeps of gtsv (tdm.c:26-48) over one strided line with the
// table's factors: q[k * st] is row k
__device__ __forceinline__ void tdm_sweeps(int m, const double * __restrict__ tab, int n,
                                           const double * __restrict__ l, double * q, size_t st) {
  const double * v = tab;
  const double * rec = tab + n;
  double prev = q[0] / rec[0];
  q[0] = prev;
  for (int i = 1; i < m - 1; i++) {
    prev = rec[i] * (q[i * st] - l[i] * prev);
    q[i * st] = prev;
  }
  if (tab[3 * n] != 0.) {
    prev = 0.;
  } else {
    prev = rec[m - 1] * (q[(m - 1) * st] - l[m - 1] * prev);
  }
  q[(m - 1) * st] = prev;
  for (int i = m - 2; i >= 0; i--) {
    prev = q[i * st] - v[i] * prev;
    q[i * st] = prev;
  }
}

// one thread: the factors (tdm.c:27,33-34,38), the periodic correction q1
// (tdm.c:80-86) and the Sherman-Morrison denominator (tdm.c:93)
__global__ void k_tdm_table(int n, int periodic, const double * __restrict__ l, const double * __restrict__ c,
                            const double * __restrict__ u, double * __restrict__ tab) {
  if (blockIdx.x != 0 || threadIdx.x != 0) return;
  const int m = periodic ? n - 1 : n;
  double * v = tab;
  double * rec = tab + n;
  v[0] = u[0] / c[0];
  rec[0] = c[0];
  for (int i = 1; i < m - 1; i++) {
    const double val = 1. / (c[i] - l[i] * v[i - 1]);
    rec[i] = val;
    v[i] = val * (u[i]);
  }
  const double val = c[m - 1] - l[m - 1] * v[m - 2];
  const bool singular = !(fabs(val) > DBL_EPSILON);
  rec[m - 1] = singular ? 0. : 1. / val;
  tab[3 * n] = singular ? 1. : 0.;
  if (!periodic) {
    tab[3 * n + 1] = 0.;
    tab[3 * n + 2] = 0.;
    return;
  }
  double * q1 = tab + 2 * n;
  for (int i = 0; i < n - 1; i++) q1[i] = i == 0 ? -l[i] : i == n - 2 ? -u[i] : 0.;
  tdm_sweeps(m, tab, n, l, q1, 1);
  const double den = c[n - 1] + u[n - 1] * q1[0] + l[n - 1] * q1[n - 2];
  tab[3 * n + 1] = den;
  tab[3 * n + 2] = fabs(den) < DBL_EPSILON ? 1. : 0.;
}

// the periodic update of one line (tdm.c:92-98)
__device__ __forceinline__ void tdm_periodic_fix(int n, const double * __restrict__ tab,
                                                 const double * __restrict__ l, const double * __restrict__ u,
                                                 double * q, size_t st) {
  const double * q1 = tab + 2 * n;
  const double num = q[(n - 1) * st] - u[n - 1] * q[0] - l[n - 1] * q[(n - 2) * st];
  const double last = tab[3 * n + 2] != 0. ? 0. : num / tab[3 * n + 1];
  q[(n - 1) * st] = last;
  for (int i = 0; i < n - 1; i++) q[i * st] = q[i * st] + last * q1[i];
}

// tdm_solve (tdm.c:66-106) over nrhs contiguous lines of n rows with ncomp
// interleaved components (1 real, 2 = the real and imaginary parts of a
// complex value, which the reference's real-by-complex arithmetic treats as
// two independent systems with the same coefficients).
//
// lpb > 0: a work-group loads lpb consecutive lines (one contiguous run of
// lpb * n * ncomp doubles, coalesced) into LDS with one system per LDS row of
// odd stride (lanes of one wavefront hit distinct banks), one lane per
// system runs the sweeps there, and the work-group stores the run back.
// lpb == 0: lines too long for LDS; one lane per system in global memory.
__global__ __launch_bounds__(256) void k_tdm_lines(int n, int nrhs, int periodic, int ncomp, int lpb,
                                                   const double * __restrict__ l, const double * __restrict__ u,
                                                   const double * __restrict__ tab, double * __restrict__ data) {
  const int m = periodic ? n - 1 : n;
  if (lpb == 0) {
    const long s = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= (long)nrhs * ncomp) return;
    double * q = data + (s / ncomp) * (size_t)n * ncomp + (s % ncomp);
    tdm_sweeps(m, tab, n, l, q, (size_t)ncomp);
    if (periodic) tdm_periodic_fix(n, tab, l, u, q, (size_t)ncomp);
    return;
  }
  extern __shared__ double lds[];
  const int stride = n | 1;
  const long j0 = (long)blockIdx.x * lpb;
  const int lines = (int)min((long)lpb, (long)nrhs - j0);
  const int run = lines * n * ncomp;
  double * g = data + j0 * (size_t)n * ncomp;
  for (int e = threadIdx.x; e < run; e += blockDim.x) {
    const int jl = e / (n * ncomp);
    const int rem = e - jl * n * ncomp;
    const int i = rem / ncomp;
    const int r = rem - i * ncomp;
    lds[(jl * ncomp + r) * stride + i] = g[e];
  }
  __syncthreads();
  const int s = threadIdx.x;
  if (s < lines * ncomp) {
    double * q = lds + s * stride;
    tdm_sweeps(m, tab, n, l, q, 1);
    if (periodic) tdm_periodic_fix(n, tab, l, u, q, 1);
  }
  __syncthreads();
  for (int e = threadIdx.x; e < run; e += blockDim.x) {
    const int jl = e / (n * ncomp);
    const int rem = e - jl * n * ncomp;
    const int i = rem / ncomp;
    const int r = rem - i * ncomp;
    g[e] = lds[(jl * ncomp + r) * stride + i];
  }
}
}  // namespace vofk

// host-only: the vtable object must not be emitted by the device pass
#if !defined(__HIP_DEVICE_COMPILE__)
// lines per work-group of k_tdm_lines within 64 KiB of LDS and 256 systems
// (0: the line does not fit, one lane per system in global memory)
static int tdm_lines_per_group(int n, int ncomp) {
  const size_t row = sizeof(double) * (size_t)(n | 1);
  const size_t by_lds = (size_t)65536 / (row * ncomp);
  const size_t by_lanes = 256 / ncomp;
  return (int)(by_lds < by_lanes ? by_lds : by_lanes);
}

// the table and line kernels of one solve on `st`
static int tdm_enqueue(int n, int nrhs, bool periodic, bool cplx, const double * l, const double * c,
                       const double * u, double * tab, void * data, hipStream_t st) {
  const int ncomp = cplx ? 2 : 1;
  hipLaunchKernelGGL(vofk::k_tdm_table, dim3(1), dim3(64), 0, st, n, (int)periodic, l, c, u, tab);
  LAUNCHCHK();
  const int lpb = tdm_lines_per_group(n, ncomp);
  if (lpb > 0) {
    const long nb = ((long)nrhs + lpb - 1) / lpb;
    const size_t shm = sizeof(double) * (size_t)(n | 1) * ncomp * lpb;
    hipLaunchKernelGGL(vofk::k_tdm_lines, dim3((unsigned)nb), dim3(256), shm, st, n, nrhs, (int)periodic, ncomp,
                       lpb, l, u, (const double *)tab, (double *)data);
  } else {
    const long nb = ((long)nrhs * ncomp + 255) / 256;
    hipLaunchKernelGGL(vofk::k_tdm_lines, dim3((unsigned)nb), dim3(256), 0, st, n, nrhs, (int)periodic, ncomp, 0,
                       l, u, (const double *)tab, (double *)data);
  }
  LAUNCHCHK();
  return 0;
}

// argument checks shared by every solve: the reference's gtsv reads v[n-2]
// (n >= 2), and its periodic path runs gtsv on n - 1 rows (n >= 3)
static int tdm_check_sizes(const char * who, int size, int nrhs, bool is_periodic) {
  if (size <= 0) return fail("ERROR(%s): size should be positive: %d", who, size);
  if (nrhs <= 0) return fail("ERROR(%s): nrhs should be positive: %d", who, nrhs);
  if (size < (is_periodic ? 3 : 2))
    return fail("ERROR(%s): size %d is too small for a%s system", who, size, is_periodic ? " periodic" : "");
  return 0;
}

static int tdm_construct(const int size, const int nrhs, const bool is_periodic, const bool is_complex,
                         tdm_info_t ** info) {
  *info = nullptr;
  if (tdm_check_sizes("construct", size, nrhs, is_periodic)) return 1;
  tdm_info_t * t = new tdm_info_t_();
  t->size = size;
  t->nrhs = nrhs;
  t->is_periodic = is_periodic;
  t->is_complex = is_complex;
  t->l = (double *)calloc(size, sizeof(double));
  t->c = (double *)calloc(size, sizeof(double));
  t->u = (double *)calloc(size, sizeof(double));
  t->d_luc = nullptr;
  t->d_tab = nullptr;
  t->stream = nullptr;
  hipError_t e = hipMalloc(&t->d_luc, sizeof(double) * 3 * size);
  if (e == hipSuccess) e = hipMalloc(&t->d_tab, sizeof(double) * vofk::tdm_table_len(size));
  if (e == hipSuccess) e = hipStreamCreate(&t->stream);
  if (e != hipSuccess || !t->l || !t->c || !t->u) {
    free(t->l); free(t->c); free(t->u);
    dev_free(t->d_luc); dev_free(t->d_tab);
    if (t->stream) (void)hipStreamDestroy(t->stream);
    delete t;
    return fail("ERROR(construct): allocation failed: %s", hipGetErrorString(e));
  }
  *info = t;
  return 0;
}
static int tdm_get_l(const tdm_info_t * info, double ** l) {
  if (!info) return fail("ERROR(get_l): info is NULL");
  *l = info->l;
  return 0;
}
static int tdm_get_c(const tdm_info_t * info, double ** c) {
  if (!info) return fail("ERROR(get_c): info is NULL");
  *c = info->c;
  return 0;
}
static int tdm_get_u(const tdm_info_t * info, double ** u) {
  if (!info) return fail("ERROR(get_u): info is NULL");
  *u = info->u;
  return 0;
}
static int tdm_get_size(const tdm_info_t * info, int * size) {
  if (!info) return fail("ERROR(get_size): info is NULL");
  *size = info->size;
  return 0;
}
static int tdm_get_nrhs(const tdm_info_t * info, int * nrhs) {
  if (!info) return fail("ERROR(get_nrhs): info is NULL");
  *nrhs = info->nrhs;
  return 0;
}
// tdm.c:269-292: synchronous, as the reference's solve returns with the answers in data
static int tdm_solve(const tdm_info_t * info, void * data) {
  if (!info) return fail("ERROR(solve): info is NULL");
  if (!data) return fail("ERROR(solve): data is NULL");
  const int n = info->size;
  hipStream_t st = info->stream;
  HIPCHK(hipMemcpyAsync(info->d_luc, info->l, sizeof(double) * n, hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(info->d_luc + n, info->c, sizeof(double) * n, hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(info->d_luc + 2 * n, info->u, sizeof(double) * n, hipMemcpyHostToDevice, st));
  if (tdm_enqueue(n, info->nrhs, info->is_periodic, info->is_complex, info->d_luc, info->d_luc + n,
                  info->d_luc + 2 * n, info->d_tab, data, st))
    return 1;
  HIPCHK(hipStreamSynchronize(st));
  return 0;
}
static int tdm_destruct(tdm_info_t * info) {
  if (!info) return fail("ERROR(destruct): info is NULL");
  free(info->l); free(info->c); free(info->u);
  dev_free(info->d_luc); dev_free(info->d_tab);
  if (info->stream) (void)hipStreamDestroy(info->stream);
  delete info;
  return 0;
}

extern "C" const tdm_t tdm = {
  tdm_construct, tdm_get_l, tdm_get_c, tdm_get_u, tdm_get_size, tdm_get_nrhs, tdm_solve, tdm_destruct,
};

// tdm_solve_noalloc: the solve of the vtable for callers that cannot
// allocate device memory (code under HIP graph capture, serving loops,
// harnesses that forbid hipMalloc).  The factor table lives in `scratch`, a
// fixed device array of this library; each GPU holds its own copy, found
// with hipGetSymbolAddress once per device (in host code the array's name is
// not a device address).  Every call on one GPU shares that GPU's copy, so
// calls that overlap on one GPU race; the table is fully rewritten (by
// k_tdm_table) before the line kernel of the same call reads it.
#endif  // !__HIP_DEVICE_COMPILE__

constexpr int kTdmNoallocMax = 16384;   // largest system size of tdm_solve_noalloc
__device__ __align__(256) double scratch[3 * kTdmNoallocMax + 3];

#if !defined(__HIP_DEVICE_COMPILE__)
static std::mutex g_scratch_mu;
static double * g_scratch_ptr[64];      // per device ordinal; null until looked up

static int scratch_for_current_device(double ** out) {
  int dev = 0;
  HIPCHK(hipGetDevice(&dev));
  if (dev < 0 || dev >= 64) return fail("tdm_solve_noalloc: device ordinal %d out of range", dev);
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void * p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return fail("tdm_solve_noalloc: hipGetSymbolAddress(scratch) failed: %s", hipGetErrorString(e));
    g_scratch_ptr[dev] = (double *)p;
  }
  *out = g_scratch_ptr[dev];
  return 0;
}

extern "C" int tdm_solve_noalloc(int size, int nrhs, bool is_periodic, bool is_complex, const double * l,
                                 const double * c, const double * u, void * data, void * stream) {
  if (tdm_check_sizes("tdm_solve_noalloc", size, nrhs, is_periodic)) return 1;
  if (size > kTdmNoallocMax)
    return fail("ERROR(tdm_solve_noalloc): size %d exceeds the fixed table's maximum %d", size, kTdmNoallocMax);
  if (!l || !c || !u || !data) return fail("ERROR(tdm_solve_noalloc): null pointer");
  double * tab = nullptr;
  if (scratch_for_current_device(&tab)) return 1;
  return tdm_enqueue(size, nrhs, is_periodic, is_complex, l, c, u, tab, data, (hipStream_t)stream);
}

#endif  // !__HIP_DEVICE_COMPILE__

#if !defined(__HIP_DEVICE_COMPILE__)
// linear_system.c:28-85 (implicit x only)
extern "C" int linear_system_init(const void * info, const bool implicit[NDIMS],
                                  const size_t glsizes[NDIMS], linear_system_t * ls) {
  (void)info;
  if (ls->is_initialised) return fail("this linear_system object is already initialised");
  for (int d = 1; d < NDIMS; d++)
    if (implicit[d]) return fail("linear_system_init: implicit treatment in direction %d is not provided", d);
  memcpy(ls->implicit, implicit, sizeof(bool) * NDIMS);
  size_t n = 1;
  for (int d = 0; d < NDIMS; d++) {
    ls->x1pncl_mysizes[d] = glsizes[d];
    ls->y1pncl_mysizes[d] = glsizes[d];
    n *= glsizes[d];
  }
  ls->x1pncl = nullptr;
  HIPCHK(hipMalloc(&ls->x1pncl, sizeof(double) * n));
  HIPCHK(hipMemset(ls->x1pncl, 0, sizeof(double) * n));
  ls->y1pncl = nullptr;
  ls->tdm_y = nullptr;
  ls->transposer_x1_to_y1 = ls->transposer_y1_to_x1 = nullptr;
  ls->tdm_x = nullptr;
  if (implicit[0]) {
    if (tdm_construct((int)glsizes[0], (int)(n / glsizes[0]), false, false, &ls->tdm_x)) return 1;
  }
  ls->is_initialised = true;
  return 0;
}

extern "C" int linear_system_finalise(linear_system_t * ls) {
  if (!ls || !ls->is_initialised) return 1;
  dev_free(ls->x1pncl);
  if (ls->implicit[0]) tdm_destruct(ls->tdm_x);
  ls->is_initialised = false;
  return 0;
}
#endif  // !__HIP_DEVICE_COMPILE__
