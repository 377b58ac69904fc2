// C ABI: plan creation/destruction, FMM matvec entry points, host helpers.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <string>
#include <vector>
#include <tuple>
#include <mutex>
#include <thread>
#include <atomic>

#include "ed_plan.h"

namespace ed {

static thread_local std::string g_last_error;

void set_error(const std::string& msg) { g_last_error = msg; }
int fail(int code, const std::string& msg) {
  g_last_error = msg;
  return code;
}

int m2l_shape(int p, int prec, int* ncp, int* kp, int* cols, int* enabled);
void gmres_free(ed_plan* P);
void gmres_invalidate_graph(ed_plan* P);

#define ED_HIP(call)                                                                   \
  do {                                                                                 \
    hipError_t e_ = (call);                                                            \
    if (e_ != hipSuccess)                                                              \
      return fail(e_ == hipErrorOutOfMemory ? ED_ERR_ALLOC : ED_ERR_DEVICE,            \
                  std::string(#call) + ": " + hipGetErrorString(e_));                  \
  } while (0)

// Round-to-nearest-even conversions of the M2L tables to 16-bit formats.
static uint16_t f32_to_bf16(float f) {
  uint32_t u;
  std::memcpy(&u, &f, 4);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}
static uint16_t f32_to_f16(float f) {
  _Float16 h = (_Float16)f;
  uint16_t u;
  std::memcpy(&u, &h, 2);
  return u;
}
static double f16_to_f64(uint16_t u) {
  _Float16 h;
  std::memcpy(&h, &u, 2);
  return (double)(float)h;
}

#define ED_TRY(x)              \
  do {                         \
    int rc_ = (x);             \
    if (rc_ != ED_OK) return rc_; \
  } while (0)

// Plan creation sends its host arrays through pinned staging memory with
// asynchronous copies on a non-blocking upload stream and synchronises once
// at the end (stage_finish), instead of one synchronous pageable copy per
// array: the host goes on building the schedules while the copies run.  The
// pinned chunks are kept for the process (one set, guarded by a mutex held
// from a plan's first staged copy to its synchronisation).
struct PinnedStage {
  std::mutex mu;
  std::vector<std::pair<char*, size_t>> chunks;
  std::map<int, hipStream_t> streams;
};
static PinnedStage& pinned_stage() {
  static PinnedStage* S = new PinnedStage();  // never destroyed: outlives every plan
  return *S;
}
struct UploadStage {
  std::unique_lock<std::mutex> lock;
  hipStream_t st = nullptr;
  size_t chunk = 0, off = 0;
  bool used = false;
};
static thread_local UploadStage* t_stage = nullptr;

// takes the staging memory and the device's upload stream on first use
static int stage_begin(UploadStage& U) {
  PinnedStage& S = pinned_stage();
  if (!U.used) {
    U.lock = std::unique_lock<std::mutex>(S.mu);
    int dev = 0;
    ED_HIP(hipGetDevice(&dev));
    auto it = S.streams.find(dev);
    if (it == S.streams.end()) {
      hipStream_t st = nullptr;
      ED_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
      it = S.streams.emplace(dev, st).first;
    }
    U.st = it->second;
    U.chunk = 0;
    U.off = 0;
    U.used = true;
  }
  return ED_OK;
}

static int h2d(void* dst, const void* src, size_t bytes) {
  if (bytes == 0) return ED_OK;
  if (!t_stage) {
    ED_HIP(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
    return ED_OK;
  }
  UploadStage& U = *t_stage;
  PinnedStage& S = pinned_stage();
  ED_TRY(stage_begin(U));
  const size_t need = (bytes + 255) & ~(size_t)255;
  while (U.chunk < S.chunks.size() && U.off + need > S.chunks[U.chunk].second) {
    ++U.chunk;
    U.off = 0;
  }
  if (U.chunk == S.chunks.size()) {
    const size_t sz = std::max<size_t>(need, (size_t)16 << 20);
    char* b = nullptr;
    ED_HIP(hipHostMalloc((void**)&b, sz, hipHostMallocDefault));
    S.chunks.emplace_back(b, sz);
  }
  char* stage = S.chunks[U.chunk].first + U.off;
  std::memcpy(stage, src, bytes);
  U.off += need;
  ED_HIP(hipMemcpyAsync(dst, stage, bytes, hipMemcpyHostToDevice, U.st));
  return ED_OK;
}

// zeros to device memory (on the upload stream while staging)
static int d_zero(void* dst, size_t bytes) {
  if (t_stage) {
    ED_TRY(stage_begin(*t_stage));
    ED_HIP(hipMemsetAsync(dst, 0, bytes, t_stage->st));
    return ED_OK;
  }
  ED_HIP(hipMemset(dst, 0, bytes));
  return ED_OK;
}

static int stage_finish(UploadStage& U) {
  int rc = ED_OK;
  if (U.used) {
    const hipError_t e = hipStreamSynchronize(U.st);
    if (e != hipSuccess) rc = fail(ED_ERR_DEVICE, std::string("plan upload: ") + hipGetErrorString(e));
    U.used = false;
    U.lock.unlock();
  }
  return rc;
}

template <typename T>
static int upload(T** dst, const std::vector<T>& src) {
  *dst = nullptr;
  const size_t bytes = std::max<size_t>(src.size(), 1) * sizeof(T);
  ED_HIP(hipMalloc((void**)dst, bytes));
  return h2d(*dst, src.data(), src.size() * sizeof(T));
}


static void sync_release(ed_plan* P);

void plan_free(ed_plan* P) {
  void* ptrs[] = {P->pts, P->area_t, P->sig_t, P->selfc_t, P->near_t, P->perm, P->node_ctr, P->node_rng,
                  P->node_parent, P->node_octant, P->first_child, P->n_child, P->M, P->L,
                  P->leaves, P->pt_node, P->p2p_items, P->p2p_seg, P->p2p_ref, P->p2s_items, P->p2s_seg, P->p2s_sym, P->p2s_tseg, P->p2s_part, P->p2s_leaf, P->ym2m, P->u_ptr, P->u_src, P->parents_list, P->v_ptr, P->v_src, P->v_tab,
                  P->m2l_pair_src, P->m2l_pair_sgn, P->m2l_pair_tgt, P->m2l_tile_tab, P->m2l_tile_beg,
                  P->m2l_tile_cnt, P->m2l_tgt_ptr, P->m2l_tgt_row, P->ypair,
                  P->near_ptr, P->near_col, P->near_w, P->cent64_t, P->area64_t, P->near_w64,
                  P->m2l_mmax, P->coarse_z, P->coarse_az, P->coarse_einv, P->coarse_cy,
                  P->coarse_part, P->coarse_tmp};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
  // t_m2l*, t_m2m, t_l2l, m2l_ascale / m2l_binv point into the shared
  // per-device operator tables
  multi_free(P);
  if (P->dense_sym) {
    dense_sym_free(*P->dense_sym);
    delete P->dense_sym;
    P->dense_sym = nullptr;
  }
  if (P->ev_fork) (void)hipEventDestroy(P->ev_fork);
  if (P->ev_join) (void)hipEventDestroy(P->ev_join);
  sync_release(P);
  sai_free(P);
  gmres_free(P);
}

// M2L translation tables depend only on the order, the lattice offset and
// the storage precision, never on the geometry: they are computed once per
// process (FP64 on the host) and reused by every later plan.
struct OperatorTables {
  std::vector<float> valu;     // nc x nc, column-major (VALU path)
  std::vector<float> f32;      // ncp x kp scaled table (FP32 MFMA)
  std::vector<uint16_t> h16;   // ncp x kp scaled table (BF16 / FP16 MFMA); for
                               // the split form, hi then lo (2 x ncp x kp FP16)
};

// The tables of one offset are computed outside the cache lock, so several
// threads can fill the cache at once; a table two threads computed together
// is kept once (both results are the same).
static const OperatorTables& cached_m2l_tables(int p, int prec, int ox, int oy, int oz, int ncp,
                                               int kp, const std::vector<double>& asc) {
  static std::mutex mu;
  static std::map<std::tuple<int, int, int, int, int>, OperatorTables> cache;
  const auto key = std::make_tuple(p, prec, ox, oy, oz);
  {
    std::lock_guard<std::mutex> lock(mu);
    auto it = cache.find(key);
    if (it != cache.end()) return it->second;
  }
  const int nc = (p + 1) * (p + 1);
  std::vector<double> Td((size_t)nc * nc);
  m2l_matrix(p, ox, oy, oz, Td.data());
  OperatorTables ot;
  if (prec < 0) {
    ot.valu.assign((size_t)nc * nc, 0.0f);
    for (int k = 0; k < nc; ++k)
      for (int j = 0; j < nc; ++j) ot.valu[(size_t)j * nc + k] = (float)Td[(size_t)k * nc + j];
  } else {
    if (prec == 0) ot.f32.assign((size_t)ncp * kp, 0.0f);
    else ot.h16.assign((size_t)ncp * kp * (prec == 3 ? 2 : 1), 0);
    for (int k = 0; k < nc; ++k)
      for (int j = 0; j < nc; ++j) {
        const double sc = Td[(size_t)k * nc + j] / (asc[k] * asc[j]);
        const size_t at = (size_t)k * kp + j;
        if (prec == 0) {
          ot.f32[at] = (float)sc;
        } else if (prec == 3) {  // sc = hi + lo to ~2^-22
          const uint16_t hi = f32_to_f16((float)sc);
          ot.h16[at] = hi;
          ot.h16[(size_t)ncp * kp + at] = f32_to_f16((float)(sc - f16_to_f64(hi)));
        } else {
          ot.h16[at] = prec == 1 ? f32_to_bf16((float)sc) : f32_to_f16((float)sc);
        }
      }
  }
  std::lock_guard<std::mutex> lock(mu);
  return cache.emplace(key, std::move(ot)).first->second;
}

// Device copies of the translation operators, shared by every plan with the
// same (device, order, M2L form, ws): the M2L table of every lattice offset a
// V list can hold (ws < max |o_i| <= 2 ws + 1, in slot order), the eight M2M
// and L2L matrices and the balanced scaling a_n^m = sqrt((n+m)!(n-m)!).
// Uploaded on first use and kept for the life of the process (20 MB for the
// order-10 split-FP16 set), so a new plan uploads no operator.  Plans point
// into it and never free it.
struct DeviceTables {
  // For the GEMM forms only offsets with non-negative components get a
  // table: a reflection R of the lattice axes maps T(o) to T(R o) =
  // D T(o) D with D = diag(+-1) (the real solid-harmonic coefficients are
  // even or odd under each axis reflection), so 56 tables (ws = 1) serve all
  // 316 offsets and each pair carries the sign pattern of its reflection.
  float* valu = nullptr;        // n_tab x nc x nc (VALU path)
  float* f32 = nullptr;         // n_tab x ncp x kp (FP32 MFMA)
  uint16_t* h16 = nullptr;      // n_tab x hsz (16-bit and split-FP16 MFMA)
  float* m2m = nullptr;         // 8 x nc x nc, column-major
  float* m2m7 = nullptr;        // octant 7's M2M alone, padded to a multiple of 4
                                // floats (16-byte aligned for float4 staging)
  float* l2l = nullptr;         // 8 x nc x nc, column-major
  float* ascale = nullptr;      // nc
  std::vector<int32_t> slot_tab;  // offset slot -> table index, -1 if none
  std::vector<uint8_t> slot_sgn;  // offset slot -> reflection (bit a: axis a negated)
  uint8_t* flip = nullptr;        // per coefficient (ncp): bit a set if it changes sign
                                  // under the reflection of axis a
  int n_tab = 0;
};

// Sign pattern of the packed real coefficients under axis reflections
// (x -> -x: cos part (-1)^m, sin part (-1)^(m+1); y -> -y: sin part -1;
// z -> -z: (-1)^(n+m)), bit 0/1/2 for x/y/z.  Both multipoles (packed
// conj(R)) and local expansions transform with this pattern.
static std::vector<uint8_t> reflection_flips(int p, int len) {
  std::vector<uint8_t> f((size_t)len, 0);
  for (int n = 0; n <= p; ++n)
    for (int m = 0; m <= n; ++m) {
      const uint8_t z = (uint8_t)(((n + m) & 1) << 2);
      if (m == 0) {
        f[(size_t)n * n] = z;
      } else {
        f[(size_t)(n * n + 2 * m - 1)] = (uint8_t)(((m & 1) ? 1 : 0) | z);
        f[(size_t)(n * n + 2 * m)] = (uint8_t)(((m & 1) ? 0 : 1) | 2 | z);
      }
    }
  return f;
}

static int device_tables(int device, int p, int kind, int ws, int ncp, int kp,
                         const DeviceTables** out) {
  static std::mutex mu;
  static std::map<std::tuple<int, int, int, int>, DeviceTables> cache;
  std::lock_guard<std::mutex> lock(mu);
  // shared tables outlive the plan being created: copied at once, not staged
  struct NoStage {
    UploadStage* saved = t_stage;
    NoStage() { t_stage = nullptr; }
    ~NoStage() { t_stage = saved; }
  } no_stage;
  const auto key = std::make_tuple(device, p, kind, ws);
  auto it = cache.find(key);
  if (it != cache.end()) {
    *out = &it->second;
    return ED_OK;
  }
  const int nc = (p + 1) * (p + 1);
  const int R = offset_radius(ws), W = 2 * R + 1;
  DeviceTables D;
  D.slot_tab.assign((size_t)W * W * W, -1);
  D.slot_sgn.assign((size_t)W * W * W, 0);
  const bool reflect = kind >= 0;  // GEMM forms: canonical tables + signs
  std::vector<int32_t> canon_tab((size_t)W * W * W, -1);
  for (int oz = -R; oz <= R; ++oz)
    for (int oy = -R; oy <= R; ++oy)
      for (int ox = -R; ox <= R; ++ox)
        if (std::max(std::abs(ox), std::max(std::abs(oy), std::abs(oz))) > ws) {
          const size_t slot = (size_t)offset_slot(ws, ox, oy, oz);
          if (!reflect) {
            D.slot_tab[slot] = D.n_tab++;
            continue;
          }
          const size_t cs = (size_t)offset_slot(ws, std::abs(ox), std::abs(oy), std::abs(oz));
          if (canon_tab[cs] < 0) canon_tab[cs] = D.n_tab++;
          D.slot_tab[slot] = canon_tab[cs];
          D.slot_sgn[slot] = (uint8_t)((ox < 0 ? 1 : 0) | (oy < 0 ? 2 : 0) | (oz < 0 ? 4 : 0));
        }
  std::vector<double> asc(nc);
  for (int n2 = 0; n2 <= p; ++n2)
    for (int m2 = 0; m2 <= n2; ++m2) {
      const double f = std::sqrt(std::tgamma(n2 + m2 + 1.0) * std::tgamma(n2 - m2 + 1.0));
      asc[n2 * n2 + (m2 == 0 ? 0 : 2 * m2 - 1)] = f;
      if (m2 > 0) asc[n2 * n2 + 2 * m2] = f;
    }
  const size_t hsz = (size_t)ncp * kp * (kind == 3 ? 2 : 1);
  std::vector<float> tv(kind < 0 ? (size_t)D.n_tab * nc * nc : 0);
  std::vector<float> tm(kind == 0 ? (size_t)D.n_tab * ncp * kp : 0);
  std::vector<uint16_t> th(kind > 0 ? (size_t)D.n_tab * hsz : 0);
  {
    // every table from its canonical offset, computed on worker threads
    std::vector<size_t> todo;
    for (size_t slot = 0; slot < D.slot_tab.size(); ++slot)
      if (D.slot_tab[slot] >= 0 && D.slot_sgn[slot] == 0) todo.push_back(slot);
    std::atomic<size_t> next{0};
    auto work = [&]() {
      for (size_t k = next++; k < todo.size(); k = next++) {
        const size_t slot = todo[k];
        const int ox = (int)(slot % W) - R, oy = (int)((slot / W) % W) - R,
                  oz = (int)(slot / (W * W)) - R;
        (void)cached_m2l_tables(p, kind, ox, oy, oz, ncp, kp, asc);
      }
    };
    const int hw = (int)std::max(1u, std::thread::hardware_concurrency());
    const int nthr = (int)std::min<size_t>((size_t)std::min(hw, 8), std::max<size_t>(todo.size(), 1));
    std::vector<std::thread> pool;
    for (int k = 1; k < nthr; ++k) pool.emplace_back(work);
    work();
    for (auto& th_ : pool) th_.join();
  }
  for (size_t slot = 0; slot < D.slot_tab.size(); ++slot) {
    const int tab = D.slot_tab[slot];
    if (tab < 0 || D.slot_sgn[slot] != 0) continue;  // each table from its canonical offset
    const int ox = (int)(slot % W) - R, oy = (int)((slot / W) % W) - R, oz = (int)(slot / (W * W)) - R;
    const OperatorTables& ot = cached_m2l_tables(p, kind, ox, oy, oz, ncp, kp, asc);
    if (kind < 0) std::memcpy(tv.data() + (size_t)tab * nc * nc, ot.valu.data(), sizeof(float) * nc * nc);
    else if (kind == 0)
      std::memcpy(tm.data() + (size_t)tab * ncp * kp, ot.f32.data(), sizeof(float) * ncp * kp);
    else
      std::memcpy(th.data() + (size_t)tab * hsz, ot.h16.data(), sizeof(uint16_t) * hsz);
  }
  // the eight M2M and eight L2L matrices, one octant and kind per task
  std::vector<float> t8((size_t)8 * nc * nc), l8((size_t)8 * nc * nc);
  {
    auto task = [&](int q) {
      std::vector<double> Td((size_t)nc * nc);
      const int o = q & 7;
      std::vector<float>& dst = q < 8 ? t8 : l8;
      if (q < 8) m2m_matrix(p, o, Td.data());
      else l2l_matrix(p, o, Td.data());
      for (int k = 0; k < nc; ++k)
        for (int j = 0; j < nc; ++j) dst[(size_t)o * nc * nc + (size_t)j * nc + k] = (float)Td[(size_t)k * nc + j];
    };
    std::vector<std::thread> pool;
    for (int q = 1; q < 16; ++q) pool.emplace_back(task, q);
    task(0);
    for (auto& th_ : pool) th_.join();
  }
  std::vector<float> af(nc);
  for (int k = 0; k < nc; ++k) af[k] = (float)asc[k];
  int rc = ED_OK;
  if (!tv.empty()) rc = upload(&D.valu, tv);
  if (rc == ED_OK && !tm.empty()) rc = upload(&D.f32, tm);
  if (rc == ED_OK && !th.empty()) rc = upload(&D.h16, th);
  if (rc == ED_OK) rc = upload(&D.m2m, t8);
  if (rc == ED_OK) {
    std::vector<float> m7((((size_t)nc * nc + 3) / 4) * 4, 0.0f);
    std::memcpy(m7.data(), t8.data() + (size_t)7 * nc * nc, sizeof(float) * nc * nc);
    rc = upload(&D.m2m7, m7);
  }
  if (rc == ED_OK) rc = upload(&D.l2l, l8);
  if (rc == ED_OK) rc = upload(&D.ascale, af);
  if (rc == ED_OK) rc = upload(&D.flip, reflection_flips(p, std::max(ncp, kp)));
  if (rc != ED_OK) {
    for (void* q : {(void*)D.valu, (void*)D.f32, (void*)D.h16, (void*)D.m2m, (void*)D.m2m7,
                    (void*)D.l2l,
                    (void*)D.ascale, (void*)D.flip})
      if (q) (void)hipFree(q);
    return rc;
  }
  *out = &cache.emplace(key, std::move(D)).first->second;
  return ED_OK;
}

// One side stream per device for the P2P branch of every plan on it, created
// on first use and kept for the life of the process (a stream holds a
// hardware queue; plans come and go, e.g. in the autotuner).
static int side_stream(int device, hipStream_t* out) {
  static std::mutex mu;
  static std::map<int, hipStream_t> streams;
  std::lock_guard<std::mutex> lock(mu);
  auto it = streams.find(device);
  if (it == streams.end()) {
    hipStream_t s = nullptr;
    ED_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    it = streams.emplace(device, s).first;
  }
  *out = it->second;
  return ED_OK;
}

// Join signals of the matvec (hipMallocSignalMemory, 8 bytes each), kept per
// device for the life of the process: a destroyed plan hands its signal, with
// its last sequence number, to the next plan on the device, so no signal
// memory is freed and a signal's value never goes down (all its writes come
// from the device's one side stream, in order).
struct SyncSignal {
  uint64_t* join = nullptr;
  uint64_t seq = 0;
};
static std::mutex g_sync_mu;
static std::map<int, std::vector<SyncSignal>> g_sync_pool;

static int sync_acquire(ed_plan* P) {
  {
    std::lock_guard<std::mutex> lock(g_sync_mu);
    auto& v = g_sync_pool[P->device];
    if (!v.empty()) {
      P->sync_join = v.back().join;
      P->sync_seq = v.back().seq;
      v.pop_back();
      return ED_OK;
    }
  }
  int can = 0;
  if (hipDeviceGetAttribute(&can, hipDeviceAttributeCanUseStreamWaitValue, P->device) !=
          hipSuccess || !can) {
    (void)hipGetLastError();
    return ED_OK;  // the join uses events
  }
  uint64_t* j = nullptr;
  if (hipExtMallocWithFlags((void**)&j, 8, hipMallocSignalMemory) != hipSuccess) {
    (void)hipGetLastError();
    return ED_OK;
  }
  ED_HIP(hipStreamWriteValue64(P->side, j, 0, 0));
  ED_HIP(hipStreamSynchronize(P->side));
  P->sync_join = j;
  P->sync_seq = 0;
  return ED_OK;
}

static void sync_release(ed_plan* P) {
  if (!P->sync_join) return;
  std::lock_guard<std::mutex> lock(g_sync_mu);
  g_sync_pool[P->device].push_back(SyncSignal{P->sync_join, P->sync_seq});
  P->sync_join = nullptr;
}

// P2P schedule: the interactions of every 64-target chunk of a leaf are cut
// into items of about kP2PItemSteps wave steps (kP2PItemSteps * slices source
// points), so that the launch holds many equal pieces of work instead of one
// block per leaf whose cost varies 60-fold with leaf and U-list size.
static int p2p_schedule(ed_plan* P, const std::vector<int32_t>& leaves) {
  const auto& H = P->host;
  const int steps = ed::kP2PItemSteps;
  std::vector<int2> seg;
  std::vector<ed::P2PItem> items;
  std::vector<int2> ref((size_t)std::max<int64_t>(H.n, 1), make_int2(0, 0));
  for (int32_t li : leaves) {
    const int32_t b = (int32_t)H.begin[li], e = (int32_t)H.end[li];
    const int32_t k0 = (int32_t)seg.size();
    int64_t total = e - b;
    seg.push_back(make_int2(b, e));
    for (int64_t q = H.u_ptr[li]; q < H.u_ptr[li + 1]; ++q) {
      const int64_t s = H.u_src[q];
      if (s == li || H.end[s] <= H.begin[s]) continue;
      seg.push_back(make_int2((int32_t)H.begin[s], (int32_t)H.end[s]));
      total += H.end[s] - H.begin[s];
    }
    const int32_t k1 = (int32_t)seg.size();
    for (int32_t t0 = b; t0 < e; t0 += 64) {
      const int32_t ntc = std::min<int32_t>(64, e - t0);
      const int32_t S = std::min<int32_t>(64 / ntc, 8);
      const int64_t budget = (int64_t)steps * S;
      const int64_t ib = (int64_t)items.size();
      int32_t k = k0, off = 0;
      int64_t left = total;
      while (left > 0) {
        const int32_t cnt = (int32_t)std::min<int64_t>(budget, left);
        ed::P2PItem it{t0, ntc, k, off, cnt, (int32_t)(items.size() * 64), S, 0};
        items.push_back(it);
        int32_t c = cnt;
        while (c > 0 && k < k1) {
          const int32_t len = seg[k].y - seg[k].x - off;
          const int32_t take = std::min(len, c);
          off += take;
          c -= take;
          if (off == seg[k].y - seg[k].x) { ++k; off = 0; }
        }
        left -= cnt;
      }
      if ((int64_t)items.size() * 64 > INT32_MAX) return fail(ED_ERR_ARG, "too many P2P items");
      for (int32_t t = 0; t < ntc; ++t)
        ref[(size_t)(t0 + t)] = make_int2((int32_t)(ib * 64 + t), (int32_t)(items.size() - ib));
    }
  }
  P->n_p2p_items = (int64_t)items.size();
  ED_TRY(upload(&P->p2p_seg, seg));
  ED_TRY(upload(&P->p2p_ref, ref));
  ED_HIP(hipMalloc(&P->p2p_items, sizeof(ed::P2PItem) * std::max<size_t>(items.size(), 1)));
  ED_TRY(h2d(P->p2p_items, items.data(), sizeof(ed::P2PItem) * items.size()));
  // (partial rows of this schedule live in the multi-RHS buffers)
  return ED_OK;
}

// One-right-hand-side P2P schedule (see ed_plan.h).  Each unordered pair of
// distinct U-list leaves goes to the larger leaf S (lower node id on a tie)
// as a source leaf; S's partner leaves are its targets, packed 64 panels to
// a chunk in pair order, and every (chunk, piece of S) is one symmetric item.
// A leaf's own interactions (and any pair listed in one direction only) run
// one-directional items like p2p_schedule's over the leaf's stream.  Rows:
// every piece of S gives each of its target leaves one row, every chunk of
// S's targets gives S one row, every piece of a leaf's one-directional
// stream one row (written chunk by chunk).
static int p2p_schedule_sym(ed_plan* P, const std::vector<int32_t>& leaves) {
  const auto& H = P->host;
  const int64_t M = (int64_t)H.level.size();
  const size_t NL = leaves.size();
  std::vector<int32_t> ord((size_t)M, -1);
  for (size_t a = 0; a < NL; ++a) ord[(size_t)leaves[a]] = (int32_t)a;
  auto nb = [&](size_t a) { return (int32_t)H.begin[leaves[a]]; };
  auto npt = [&](size_t a) { return (int32_t)(H.end[leaves[a]] - H.begin[leaves[a]]); };
  // directed U pairs (target, source) in leaf ordinals, sorted
  std::vector<std::pair<int32_t, int32_t>> dir;
  dir.reserve((size_t)H.u_src.size());
  for (size_t a = 0; a < NL; ++a) {
    const int32_t li = leaves[a];
    for (int64_t q = H.u_ptr[li]; q < H.u_ptr[li + 1]; ++q)
      if (H.u_src[q] != li) dir.emplace_back((int32_t)a, ord[(size_t)H.u_src[q]]);
  }
  std::sort(dir.begin(), dir.end());
  std::vector<int32_t> dptr(NL + 1, 0);  // dir entries of target a: [dptr[a], dptr[a + 1])
  for (const auto& d : dir) ++dptr[(size_t)d.first + 1];
  for (size_t a = 0; a < NL; ++a) dptr[a + 1] += dptr[a];
  auto listed = [&](int32_t t, int32_t s_) {  // s_ in the U list of t
    return std::binary_search(dir.begin() + dptr[(size_t)t], dir.begin() + dptr[(size_t)t + 1],
                              std::make_pair(t, s_));
  };
  // partners of each source leaf (CSR) and the one-directional extra
  // sources; pairs listed in one direction only go to the target's stream
  std::vector<int32_t> pcnt(NL + 1, 0), xcnt(NL + 1, 0);
  std::vector<uint8_t> kind(dir.size(), 0);  // 0 skip, 1 symmetric (A source), 2 (B source), 3 one-way
  for (size_t e = 0; e < dir.size(); ++e) {
    const int32_t A = dir[e].first, B = dir[e].second;
    if (npt((size_t)A) == 0 || npt((size_t)B) == 0) continue;
    if (!listed(B, A)) {
      kind[e] = 3;
      ++xcnt[(size_t)A + 1];
    } else if (A < B) {
      const bool a_src = npt((size_t)A) >= npt((size_t)B);
      kind[e] = a_src ? 1 : 2;
      ++pcnt[(size_t)(a_src ? A : B) + 1];
    }
  }
  for (size_t a = 0; a < NL; ++a) { pcnt[a + 1] += pcnt[a]; xcnt[a + 1] += xcnt[a]; }
  std::vector<int32_t> partner((size_t)pcnt[NL]), extra((size_t)xcnt[NL]);
  {
    std::vector<int32_t> pf(pcnt.begin(), pcnt.end() - 1), xf(xcnt.begin(), xcnt.end() - 1);
    for (size_t e = 0; e < dir.size(); ++e) {
      const int32_t A = dir[e].first, B = dir[e].second;
      if (kind[e] == 1) partner[(size_t)pf[(size_t)A]++] = B;
      else if (kind[e] == 2) partner[(size_t)pf[(size_t)B]++] = A;
      else if (kind[e] == 3) extra[(size_t)xf[(size_t)A]++] = B;
    }
  }
  // pass 1: rows of every leaf and item counts
  constexpr int32_t PS = ed::kP2PSymSources;
  const int steps = ed::kP2PItemSteps;
  std::vector<int32_t> rows(NL, 0), kT0((size_t)pcnt[NL]), kS0(NL, 0), nchunk(NL, 0),
      npc(NL, 0), k1d(NL, 0), np1d(NL, 0);
  std::vector<int64_t> budget1d(NL, 0), total1d(NL, 0);
  int64_t n_sym = 0, n_one = 0, sym_int = 0;
  for (size_t S = 0; S < NL; ++S) {
    if (pcnt[S + 1] == pcnt[S]) continue;
    npc[S] = (npt(S) + PS - 1) / PS;
    int64_t tot = 0;
    for (int32_t e = pcnt[S]; e < pcnt[S + 1]; ++e) {
      const size_t T = (size_t)partner[(size_t)e];
      kT0[(size_t)e] = rows[T];
      rows[T] += npc[S];
      tot += npt(T);
    }
    nchunk[S] = (int32_t)((tot + 63) / 64);
    kS0[S] = rows[S];
    rows[S] += nchunk[S];
    n_sym += (int64_t)nchunk[S] * npc[S];
    sym_int += 2 * tot * npt(S);
  }
  for (size_t a = 0; a < NL; ++a) {
    const int32_t n_a = npt(a);
    if (n_a == 0) continue;
    int64_t total = n_a;
    for (int32_t e = xcnt[a]; e < xcnt[a + 1]; ++e) total += npt((size_t)extra[(size_t)e]);
    // pieces of one source count for every chunk (kP2PItemSteps per slice of
    // the leaf's first, fullest chunk), so piece q of every chunk writes row q
    const int32_t S0 = std::min<int32_t>(64 / std::min<int32_t>(64, n_a), 8);
    budget1d[a] = (int64_t)steps * S0;
    total1d[a] = total;
    np1d[a] = (int32_t)((total + budget1d[a] - 1) / budget1d[a]);
    k1d[a] = rows[a];
    rows[a] += np1d[a];
    n_one += (int64_t)((n_a + 63) / 64) * np1d[a];
  }
  std::vector<int64_t> base(NL + 1, 0);
  for (size_t a = 0; a < NL; ++a) base[a + 1] = base[a] + (int64_t)rows[a] * npt(a);
  if (base[NL] > INT32_MAX || n_sym > INT32_MAX)
    return fail(ED_ERR_ARG, "too many P2P partial slots");
  // pass 2: symmetric items; per chunk its target segments (one per target
  // leaf in the chunk), shared by the chunk's pieces
  std::vector<ed::P2PSymItem> sym((size_t)n_sym);
  std::vector<ed::P2PSymSeg> tseg;
  tseg.reserve((size_t)n_sym * 2);
  size_t it = 0;
  for (size_t S = 0; S < NL; ++S) {
    if (nchunk[S] == 0) continue;
    const int32_t sb = nb(S), nS = npt(S);
    int32_t e = pcnt[S], off = 0;  // next target: partner e, panel offset off
    for (int32_t c = 0; c < nchunk[S]; ++c) {
      const int32_t g0 = (int32_t)tseg.size();
      int32_t nl = 0;
      while (nl < 64 && e < pcnt[S + 1]) {
        const size_t T = (size_t)partner[(size_t)e];
        const int32_t nT = npt(T);
        const int32_t take = std::min(64 - nl, nT - off);
        const int64_t slot0 = base[T] + (int64_t)kT0[(size_t)e] * nT + off;
        tseg.push_back(ed::P2PSymSeg{nl, nb(T) + off, (int32_t)slot0, nT});
        nl += take;
        off += take;
        if (off == nT) { ++e; off = 0; }
      }
      const int32_t ng = (int32_t)tseg.size() - g0;
      const int64_t srow = base[S] + (int64_t)(kS0[S] + c) * nS;
      for (int32_t q = 0; q < npc[S]; ++q, ++it) {
        const int32_t s0 = sb + q * PS, cnt = std::min(PS, sb + nS - s0);
        sym[it] = ed::P2PSymItem{g0, ng, q, nl, s0, cnt, (int32_t)(srow + (s0 - sb)), 0};
      }
    }
  }
  // one-directional items: chunks of <= 64 targets over the leaf's stream
  // (itself, then its one-way sources)
  std::vector<int2> seg;
  std::vector<ed::P2PItem> items;
  items.reserve((size_t)n_one);
  for (size_t a = 0; a < NL; ++a) {
    const int32_t b = nb(a), e = b + npt(a);
    if (e <= b) continue;
    for (int32_t t0 = b; t0 < e; t0 += 64) {
      const int32_t ntc = std::min<int32_t>(64, e - t0);
      const int32_t Sl = std::min<int32_t>(64 / ntc, 8);
      const int32_t kseg0 = (int32_t)seg.size();
      seg.push_back(make_int2(b, e));
      for (int32_t x = xcnt[a]; x < xcnt[a + 1]; ++x) {
        const size_t B = (size_t)extra[(size_t)x];
        seg.push_back(make_int2(nb(B), nb(B) + npt(B)));
      }
      const int32_t kseg1 = (int32_t)seg.size();
      int32_t k = kseg0, so = 0;
      int64_t left = total1d[a];
      for (int32_t q = 0; q < np1d[a]; ++q) {
        const int32_t cnt = (int32_t)std::min<int64_t>(budget1d[a], left);
        const int64_t slot = base[a] + (int64_t)(k1d[a] + q) * (e - b) + (t0 - b);
        items.push_back(ed::P2PItem{t0, ntc, k, so, cnt, (int32_t)slot, Sl, 0});
        int32_t c = cnt;
        while (c > 0 && k < kseg1) {
          const int32_t len = seg[k].y - seg[k].x - so;
          const int32_t take = std::min(len, c);
          so += take;
          c -= take;
          if (so == seg[k].y - seg[k].x) { ++k; so = 0; }
        }
        left -= cnt;
      }
    }
  }
  std::vector<int2> leafrow((size_t)std::max<int64_t>(M, 1), make_int2(0, 0));
  for (size_t a = 0; a < NL; ++a) leafrow[(size_t)leaves[a]] = make_int2((int32_t)base[a], rows[a]);
  P->n_p2s_items = (int64_t)items.size();
  P->n_p2s_sym = (int64_t)sym.size();
  P->n_p2s_slots = base[NL];
  P->n_p2s_sym_interactions = sym_int;
  ED_TRY(upload(&P->p2s_seg, seg));
  ED_TRY(upload(&P->p2s_leaf, leafrow));
  ED_HIP(hipMalloc(&P->p2s_tseg, sizeof(ed::P2PSymSeg) * std::max<size_t>(tseg.size(), 1)));
  ED_TRY(h2d(P->p2s_tseg, tseg.data(), sizeof(ed::P2PSymSeg) * tseg.size()));
  ED_HIP(hipMalloc(&P->p2s_items, sizeof(ed::P2PItem) * std::max<size_t>(items.size(), 1)));
  ED_TRY(h2d(P->p2s_items, items.data(), sizeof(ed::P2PItem) * items.size()));
  ED_HIP(hipMalloc(&P->p2s_sym, sizeof(ed::P2PSymItem) * std::max<size_t>(sym.size(), 1)));
  ED_TRY(h2d(P->p2s_sym, sym.data(), sizeof(ed::P2PSymItem) * sym.size()));
  ED_HIP(hipMalloc((void**)&P->p2s_part, sizeof(float) * (size_t)std::max<int64_t>(base[NL], 1)));
  return ED_OK;
}

int plan_upload(ed_plan* P) {
  const HostTree& H = P->host;
  const int64_t n = H.n, M = (int64_t)H.level.size();
  const int p = P->p, nc = n_real(p);
  P->nc = nc;
  P->n = n;
  P->n_nodes = M;
  // points (relative to the root centre), perm, areas handled by caller
  std::vector<float4> pts(n);
  (void)pts;
  // nodes
  std::vector<float4> ctr(M);
  std::vector<int2> rng(M);
  std::vector<int32_t> par(M), oct(M), fc(M), nch(M);
  for (int64_t i = 0; i < M; ++i) {
    const double s = H.node_s[i];
    ctr[i] = make_float4((float)H.node_c[3 * i], (float)H.node_c[3 * i + 1],
                         (float)H.node_c[3 * i + 2], (float)(1.0 / s));
    rng[i] = make_int2((int)H.begin[i], (int)H.end[i]);
    par[i] = (int32_t)std::max<int64_t>(H.parent[i], 0);
    oct[i] = (H.coord[3 * i] & 1) | ((H.coord[3 * i + 1] & 1) << 1) | ((H.coord[3 * i + 2] & 1) << 2);
    fc[i] = H.first_child[i] < 0 ? 0 : H.first_child[i];
    nch[i] = H.n_child[i];
  }
  ED_TRY(upload(&P->node_ctr, ctr));
  ED_TRY(upload(&P->node_rng, rng));
  ED_TRY(upload(&P->node_parent, par));
  ED_TRY(upload(&P->node_octant, oct));
  ED_TRY(upload(&P->first_child, fc));
  ED_TRY(upload(&P->n_child, nch));
  ED_HIP(hipMalloc((void**)&P->M, sizeof(float) * (size_t)M * nc));
  ED_HIP(hipMalloc((void**)&P->L, sizeof(float) * (size_t)M * nc));
  ED_HIP(hipMalloc((void**)&P->ym2m, sizeof(float) * (size_t)M * nc));
  ED_TRY(d_zero(P->M, sizeof(float) * (size_t)M * nc));
  ED_TRY(d_zero(P->L, sizeof(float) * (size_t)M * nc));

  // leaves and U lists
  std::vector<int32_t> leaves, uptr{0}, usrc;
  int64_t p2p_int = 0, p2p_src = 0;
  for (int64_t i = 0; i < M; ++i) {
    if (!H.is_leaf[i]) continue;
    leaves.push_back((int32_t)i);
    for (int64_t q = H.u_ptr[i]; q < H.u_ptr[i + 1]; ++q) {
      const int64_t s = H.u_src[q];
      usrc.push_back((int32_t)s);
      p2p_int += (H.end[i] - H.begin[i]) * (H.end[s] - H.begin[s]);
      p2p_src += H.end[s] - H.begin[s];
    }
    uptr.push_back((int32_t)usrc.size());
  }
  P->n_leaves = (int64_t)leaves.size();
  P->n_u = (int64_t)usrc.size();
  P->n_p2p_interactions = p2p_int;
  P->n_p2p_src = p2p_src;
  ED_TRY(upload(&P->leaves, leaves));
  {
    std::vector<int32_t> ptn((size_t)H.n, 0);
    for (int32_t li : leaves)
      for (int64_t t = H.begin[li]; t < H.end[li]; ++t) ptn[(size_t)t] = li;
    ED_TRY(upload(&P->pt_node, ptn));
  }
  ED_TRY(p2p_schedule(P, leaves));
  ED_TRY(p2p_schedule_sym(P, leaves));
  ED_TRY(side_stream(P->device, &P->side));
  ED_HIP(hipDeviceGetAttribute(&P->n_cu, hipDeviceAttributeMultiprocessorCount, P->device));
  ED_HIP(hipEventCreateWithFlags(&P->ev_fork, hipEventDisableTiming));
  ED_HIP(hipEventCreateWithFlags(&P->ev_join, hipEventDisableTiming));
  ED_TRY(sync_acquire(P));

  ED_TRY(upload(&P->u_ptr, uptr));
  ED_TRY(upload(&P->u_src, usrc));

  // parents per level
  std::vector<int32_t> plist;
  P->parents_begin.assign(H.depth + 2, 0);
  for (int l = 0; l <= H.depth; ++l) {
    P->parents_begin[l] = (int64_t)plist.size();
    for (int64_t i = H.level_begin[l]; i < H.level_begin[l + 1]; ++i)
      if (H.n_child[i] > 0) plist.push_back((int32_t)i);
  }
  P->parents_begin[H.depth + 1] = (int64_t)plist.size();
  ED_TRY(upload(&P->parents_list, plist));

  // V lists and offset tables
  const int prec = P->m2l_precision;
  int ncp = 0, kp = 0, cols = 0, enabled = 0;
  ED_TRY(m2l_shape(p, prec < 0 ? 0 : prec, &ncp, &kp, &cols, &enabled));
  if (prec < 0) enabled = 0;
  P->ncp = ncp;
  const bool need_valu = !enabled || prec < 0;
  const DeviceTables* DT = nullptr;
  ED_TRY(device_tables(P->device, p, need_valu ? -1 : prec, P->ws, ncp, kp, &DT));
  const int ntab = DT->n_tab;
  std::vector<int32_t> vptr(M + 1), vsrc(H.v_src.size()), vtab(H.v_src.size());
  std::vector<uint8_t> vsgn(H.v_src.size());
  std::vector<char> used(DT->slot_tab.size(), 0);
  for (int64_t i = 0; i <= M; ++i) vptr[i] = (int32_t)H.v_ptr[i];
  for (size_t q = 0; q < H.v_src.size(); ++q) {
    vsrc[q] = (int32_t)H.v_src[q];
    const int32_t s = H.v_off[q];
    const int32_t tab = s >= 0 && (size_t)s < DT->slot_tab.size() ? DT->slot_tab[(size_t)s] : -1;
    if (tab < 0) return fail(ED_ERR_INTERNAL, "plan: V-list offset outside the table set");
    vtab[q] = tab;
    vsgn[q] = DT->slot_sgn[(size_t)s];
    used[(size_t)s] = 1;
  }
  P->n_off = (int)std::count(used.begin(), used.end(), (char)1);  // distinct offsets in use
  P->n_v = (int64_t)vsrc.size();
  {
    // The P2P grid cap leaves the far-field chain's kernels room beside the
    // P2P; it pays only while the chain is the longer branch.  Branch lengths
    // from rates measured on MI355X at p = 10: the chain about 1.27 ns per
    // M2L pair (82k sphere: 66 us for 52,160 pairs; 300k torus: 173 us for
    // 136,112), taken to scale with the coefficient count, and the symmetric
    // P2P about 3.8e6 evaluations per us (46 us for 157 M, 199 us for 843 M).
    const double nc = (double)((P->p + 1) * (P->p + 1));
    const double chain_us = 1.27e-3 * (nc / 121.0) * (double)P->n_v;
    const double p2p_us = (double)P->n_p2s_sym_interactions / 3.8e6;
    P->p2p_blocks_per_cu = p2p_us < chain_us ? ed::kP2PBlocksPerCU : 0;
  }
  ED_TRY(upload(&P->v_ptr, vptr));
  ED_TRY(upload(&P->v_src, vsrc));
  ED_TRY(upload(&P->v_tab, vtab));
  P->t_m2l = DT->valu;
  P->t_m2l_mfma = DT->f32;
  P->t_m2l_h = DT->h16;
  if (enabled && (prec == 1 || prec == 2)) {
    // split-FP16 tables for the solver's residual updates: same shape (the
    // 16-bit forms share ncp, kp and the tile schedule) and, built from the
    // same offsets, the same table ids
    const DeviceTables* D3 = nullptr;
    ED_TRY(device_tables(P->device, p, 3, P->ws, ncp, kp, &D3));
    if (D3->n_tab != ntab || D3->slot_tab != DT->slot_tab)
      return fail(ED_ERR_INTERNAL, "plan: split-FP16 table set differs from the 16-bit one");
    P->t_m2l_h3 = D3->h16;
  }
  P->t_m2m = DT->m2m;
  P->t_m2m7 = reinterpret_cast<const float4*>(DT->m2m7);
  P->t_l2l = DT->l2l;
  // T = diag(a) T_s diag(a): the GEMM kernels multiply each gathered
  // multipole coefficient by a (ascale) and the reduce multiplies each summed
  // row by a (binv), so L = T M exactly up to rounding.
  P->m2l_ascale = DT->ascale;
  P->m2l_binv = DT->ascale;
  P->m2l_flip = DT->flip;
  ED_TRY(upload(&P->m2l_mmax, std::vector<unsigned>(64 * 32, 0u)));  // kMaxSlots x kMaxSlotStride

  // GEMM-form M2L schedule: pairs grouped by table, tiles of `cols` pairs
  P->n_m2l_tiles = 0;
  if (enabled && P->n_v > 0) {
    std::vector<std::vector<std::pair<int32_t, int32_t>>> by_tab(ntab);
    std::vector<std::vector<uint8_t>> sgn_tab(ntab);
    for (int64_t t = 0; t < M; ++t)
      for (int64_t q = H.v_ptr[t]; q < H.v_ptr[t + 1]; ++q) {
        by_tab[vtab[q]].emplace_back((int32_t)t, vsrc[q]);
        sgn_tab[vtab[q]].push_back(vsgn[q]);
      }
    std::vector<int32_t> psrc, ptgt, ttab, tbeg, tcnt;
    std::vector<uint8_t> psgn;
    for (int tab = 0; tab < ntab; ++tab) {
      auto& v = by_tab[tab];
      for (size_t b = 0; b < v.size(); b += (size_t)cols) {
        ttab.push_back(tab);
        tbeg.push_back((int32_t)psrc.size() + 0);
        tcnt.push_back((int32_t)std::min<size_t>((size_t)cols, v.size() - b));
        for (size_t c = b; c < std::min(v.size(), b + (size_t)cols); ++c) {
          ptgt.push_back(v[c].first);
          psrc.push_back(v[c].second);
          psgn.push_back(sgn_tab[tab][c]);
        }
      }
    }
    // tiles were appended with tbeg = running pair count before each tile
    {
      int32_t run = 0;
      for (size_t t = 0; t < ttab.size(); ++t) { tbeg[t] = run; run += tcnt[t]; }
    }
    std::vector<int32_t> tptr(M + 1, 0), trow(psrc.size());
    for (int32_t t : ptgt) tptr[t + 1]++;
    for (int64_t t = 0; t < M; ++t) tptr[t + 1] += tptr[t];
    std::vector<int32_t> fillp(tptr.begin(), tptr.end() - 1);
    for (size_t q = 0; q < ptgt.size(); ++q) trow[fillp[ptgt[q]]++] = (int32_t)q;
    // XCD-aware launch order: workgroups go to the 8 XCDs round robin
    // (block b -> XCD b % 8), each XCD with its own L2.  Tables are binned to
    // XCDs (largest tile count first, into the lightest bin) and block
    // 8 k + x runs the k-th tile of bin x, so every tile of a table reads it
    // from the same L2.  Bins are padded with empty tiles (count 0).
    {
      constexpr int kXcd = 8;
      std::vector<std::vector<int32_t>> tiles_of(ntab);
      for (size_t t = 0; t < ttab.size(); ++t) tiles_of[ttab[t]].push_back((int32_t)t);
      std::vector<int32_t> order(ntab);
      for (int t = 0; t < ntab; ++t) order[t] = t;
      std::stable_sort(order.begin(), order.end(), [&](int32_t a, int32_t b) {
        return tiles_of[a].size() > tiles_of[b].size();
      });
      std::vector<std::vector<int32_t>> bins(kXcd);
      for (int32_t tab : order) {
        int best = 0;
        for (int x = 1; x < kXcd; ++x)
          if (bins[x].size() < bins[best].size()) best = x;
        for (int32_t t : tiles_of[tab]) bins[best].push_back(t);
      }
      size_t depth = 0;
      for (auto& b : bins) depth = std::max(depth, b.size());
      std::vector<int32_t> xt, xb, xc;
      for (size_t k = 0; k < depth; ++k)
        for (int x = 0; x < kXcd; ++x) {
          if (k < bins[x].size()) {
            const int32_t t = bins[x][k];
            xt.push_back(ttab[t]); xb.push_back(tbeg[t]); xc.push_back(tcnt[t]);
          } else {
            xt.push_back(0); xb.push_back(0); xc.push_back(0);
          }
        }
      ttab.swap(xt);
      tbeg.swap(xb);
      tcnt.swap(xc);
    }
    ED_TRY(upload(&P->m2l_pair_src, psrc));
    ED_TRY(upload(&P->m2l_pair_sgn, psgn));
    ED_TRY(upload(&P->m2l_pair_tgt, ptgt));
    ED_TRY(upload(&P->m2l_tile_tab, ttab));
    ED_TRY(upload(&P->m2l_tile_beg, tbeg));
    ED_TRY(upload(&P->m2l_tile_cnt, tcnt));
    ED_TRY(upload(&P->m2l_tgt_ptr, tptr));
    ED_TRY(upload(&P->m2l_tgt_row, trow));
    ED_HIP(hipMalloc((void**)&P->ypair, sizeof(float) * std::max<size_t>(psrc.size(), 1) * ncp));
    P->n_m2l_tiles = (int64_t)ttab.size();
  }

  // empty near CSR
  std::vector<int32_t> nptr(n + 1, 0);
  ED_TRY(upload(&P->near_ptr, nptr));
  ED_HIP(hipMalloc((void**)&P->near_col, sizeof(int32_t)));
  ED_HIP(hipMalloc((void**)&P->near_w, sizeof(float)));
  ED_HIP(hipMalloc((void**)&P->near_w64, sizeof(double)));
  P->n_near = 0;
  return ED_OK;
}

}  // namespace ed

namespace ed {
void multi_free(ed_plan* P) {
  auto& m = P->multi;
  for (void* q : {(void*)m.qT, (void*)m.sig, (void*)m.M, (void*)m.L, (void*)m.ym2m, (void*)m.ypair,
                  (void*)m.part, (void*)m.mmax})
    if (q) (void)hipFree(q);
  m = ed_plan::Multi();
}

int multi_reserve(ed_plan* P, int nrhs) {
  if (P->multi.cap >= nrhs) return ED_OK;
  gmres_invalidate_graph(P);  // a captured Arnoldi step refers to the old buffers
  multi_free(P);
  auto& m = P->multi;
  const size_t R = (size_t)nrhs, n = (size_t)P->n, mn = (size_t)P->n_nodes * P->nc;
  const size_t pairs = (size_t)std::max<int64_t>(P->n_v, 1) * (size_t)std::max(P->ncp, 1);
  const size_t part = 64 * (size_t)std::max<int64_t>(P->n_p2p_items, 1);
  const size_t slots = 64 * 32;  // kMaxSlots x kMaxSlotStride
  bool ok = hipMalloc((void**)&m.qT, sizeof(float) * n * R) == hipSuccess &&
            hipMalloc((void**)&m.sig, sizeof(float) * n * R) == hipSuccess &&
            hipMalloc((void**)&m.M, sizeof(float) * mn * R) == hipSuccess &&
            hipMalloc((void**)&m.L, sizeof(float) * mn * R) == hipSuccess &&
            hipMalloc((void**)&m.ym2m, sizeof(float) * mn * R) == hipSuccess &&
            hipMalloc((void**)&m.ypair, sizeof(float) * pairs * R) == hipSuccess &&
            hipMalloc((void**)&m.part, sizeof(float) * part * R) == hipSuccess &&
            hipMalloc((void**)&m.mmax, sizeof(unsigned) * slots * R) == hipSuccess;
  // nodes that no pass writes (levels 0 and 1) read as zero, as in the
  // one-RHS buffers
  ok = ok && hipMemset(m.M, 0, sizeof(float) * mn * R) == hipSuccess &&
       hipMemset(m.L, 0, sizeof(float) * mn * R) == hipSuccess &&
       hipMemset(m.mmax, 0, sizeof(unsigned) * slots * R) == hipSuccess;
  if (!ok) {
    multi_free(P);
    return fail(ED_ERR_ALLOC, "multi-RHS matvec: alloc");
  }
  m.cap = nrhs;
  m.pair_stride = (int64_t)pairs;
  return ED_OK;
}
}  // namespace ed

using namespace ed;

extern "C" {

const char* ed_last_error(void) { return g_last_error.c_str(); }
int32_t ed_abi_version(void) { return 1; }

void ed_plan_opts_default(ed_plan_opts* o) {
  if (!o) return;
  std::memset(o, 0, sizeof(*o));
  o->leaf_size = 64;
  o->order = 8;
  o->ws = 1;
  o->max_level = 16;
  o->m2l_precision = 0;
}

static int plan_create_impl(const double* centroids, const double* areas, int64_t n,
                            const int64_t* geom_off, int64_t n_geoms, const ed_plan_opts* opts,
                            ed_plan** out) {
  if (!out) return fail(ED_ERR_ARG, "ed_plan_create: out is NULL");
  *out = nullptr;
  if (!centroids || !areas) return fail(ED_ERR_ARG, "ed_plan_create: NULL geometry");
  if (n <= 0 || n >= (1LL << 31)) return fail(ED_ERR_ARG, "ed_plan_create: n out of range");
  ed_plan_opts o;
  ed_plan_opts_default(&o);
  if (opts) o = *opts;
  if (o.order != 4 && o.order != 6 && o.order != 8 && o.order != 10 && o.order != 12 &&
      o.order != 14 && o.order != 16)
    return fail(ED_ERR_ARG, "ed_plan_create: order must be one of 4, 6, 8, 10, 12, 14, 16");
  if (o.m2l_precision < -1 || o.m2l_precision > 3)
    return fail(ED_ERR_ARG, "ed_plan_create: m2l_precision must be -1, 0, 1, 2 or 3");
  if (n_geoms < 1) return fail(ED_ERR_ARG, "ed_plan_create: n_geoms must be >= 1");
  if (geom_off) {
    if (geom_off[0] != 0 || geom_off[n_geoms] != n)
      return fail(ED_ERR_ARG, "ed_plan_create_batched: offsets must run from 0 to n");
    for (int64_t g = 0; g < n_geoms; ++g)
      if (geom_off[g + 1] <= geom_off[g])
        return fail(ED_ERR_ARG, "ed_plan_create_batched: every geometry needs >= 1 panel");
  }
  for (int64_t i = 0; i < n; ++i)
    if (!std::isfinite(areas[i])) return fail(ED_ERR_ARG, "ed_plan_create: non-finite area");
  int dev = 0;
  ED_HIP(hipGetDevice(&dev));
  ed_plan* P = new ed_plan();
  P->device = dev;
  P->p = o.order;
  P->ws = o.ws;
  P->leaf_size = o.leaf_size;
  P->max_level = o.max_level;
  P->m2l_precision = o.m2l_precision;
  int rc = ED_OK;
  if (!geom_off || n_geoms == 1) {
    rc = build_host_tree(centroids, n, o.leaf_size, o.max_level, o.ws, P->host);
  } else {
    std::vector<HostTree> parts((size_t)n_geoms);
    for (int64_t g = 0; g < n_geoms && rc == ED_OK; ++g)
      rc = build_host_tree(centroids + 3 * geom_off[g], geom_off[g + 1] - geom_off[g],
                           o.leaf_size, o.max_level, o.ws, parts[(size_t)g]);
    if (rc == ED_OK) rc = merge_host_trees(parts, P->host);
  }
  UploadStage stage;
  t_stage = &stage;
  if (rc == ED_OK) rc = plan_upload(P);
  if (rc == ED_OK) {
    const HostTree& H = P->host;
    // pts[n]: a zero-charge point far outside every geometry, the source that
    // the P2P's last group of a source range reads past the range's end
    // (it adds 0; k_gather writes only pts[0 .. n-1])
    std::vector<float4> pts(n + 1);
    pts[n] = make_float4(1e18f, 1e18f, 1e18f, 0.0f);
    std::vector<float> at(n);
    std::vector<int32_t> perm(n);
    std::vector<double> c64(3 * n), a64(n);
    const int64_t G = (int64_t)H.geom_pt_off.size() - 1;
    for (int64_t g = 0; g < G; ++g) {
      const double* gc = H.geom_center.data() + 3 * g;
      for (int64_t i = H.geom_pt_off[g]; i < H.geom_pt_off[g + 1]; ++i) {
        const int64_t o2 = H.perm[i];
        pts[i] = make_float4((float)(centroids[3 * o2] - gc[0]),
                             (float)(centroids[3 * o2 + 1] - gc[1]),
                             (float)(centroids[3 * o2 + 2] - gc[2]), 0.0f);
        at[i] = (float)areas[o2];
        perm[i] = (int32_t)o2;
        for (int a = 0; a < 3; ++a) c64[3 * i + a] = centroids[3 * o2 + a];
        a64[i] = areas[o2];
      }
    }
    rc = upload(&P->pts, pts);
    if (rc == ED_OK) rc = upload(&P->cent64_t, c64);
    if (rc == ED_OK) rc = upload(&P->area64_t, a64);
    if (rc == ED_OK) rc = upload(&P->area_t, at);
    if (rc == ED_OK) rc = upload(&P->perm, perm);
    if (rc == ED_OK) {
      hipError_t e1 = hipMalloc((void**)&P->sig_t, sizeof(float) * n);
      hipError_t e2 = hipMalloc((void**)&P->selfc_t, sizeof(float) * n);
      hipError_t e3 = hipMalloc((void**)&P->near_t, sizeof(float) * n);
      if (e1 != hipSuccess || e2 != hipSuccess || e3 != hipSuccess)
        rc = fail(ED_ERR_ALLOC, "ed_plan_create: alloc");
    }
  }
  t_stage = nullptr;
  {
    const int rc2 = stage_finish(stage);  // always: releases the staging memory
    if (rc == ED_OK) rc = rc2;
  }
  if (rc != ED_OK) {
    plan_free(P);
    delete P;
    return rc;
  }
  *out = P;
  return ED_OK;
}

int ed_plan_create(const double* centroids, const double* areas, int64_t n,
                   const ed_plan_opts* opts, ed_plan** out) {
  return plan_create_impl(centroids, areas, n, nullptr, 1, opts, out);
}

int ed_plan_create_batched(const double* centroids, const double* areas, int64_t n,
                           const int64_t* geom_offsets, int64_t n_geoms,
                           const ed_plan_opts* opts, ed_plan** out) {
  if (!geom_offsets) return fail(ED_ERR_ARG, "ed_plan_create_batched: NULL offsets");
  return plan_create_impl(centroids, areas, n, geom_offsets, n_geoms, opts, out);
}

int ed_plan_geometries(const ed_plan* P, int64_t* n_geoms_out, int64_t* offsets_host,
                       int64_t capacity) {
  if (!P || !n_geoms_out) return fail(ED_ERR_ARG, "ed_plan_geometries: NULL argument");
  const int64_t G = (int64_t)P->host.geom_pt_off.size() - 1;
  *n_geoms_out = G;
  if (offsets_host) {
    if (capacity < G + 1) return fail(ED_ERR_ARG, "ed_plan_geometries: capacity");
    for (int64_t g = 0; g <= G; ++g) offsets_host[g] = P->host.geom_pt_off[g];
  }
  return ED_OK;
}

int ed_plan_destroy(ed_plan* P) {
  if (!P) return ED_OK;
  plan_free(P);
  delete P;
  return ED_OK;
}

int ed_plan_stats(const ed_plan* P, int64_t* out, int32_t n_out) {
  if (!P || !out) return fail(ED_ERR_ARG, "ed_plan_stats: NULL argument");
  int levels_with_m2l = 0;
  for (int l = 0; l <= P->host.depth; ++l) {
    bool any = false;
    for (int64_t t = P->host.level_begin[l]; t < P->host.level_begin[l + 1] && !any; ++t)
      any = P->host.v_ptr[t + 1] > P->host.v_ptr[t];
    levels_with_m2l += any;
  }
  const int64_t vals[16] = {P->n, P->n_nodes, P->n_leaves, P->host.depth, P->n_u,
                            P->n_p2p_interactions, P->n_v, P->n_off, P->p, P->ws,
                            P->n_near, levels_with_m2l, P->n_p2p_src,
                            P->n_p2s_sym_interactions, P->n_p2s_sym + P->n_p2s_items,
                            P->n_p2s_slots};
  for (int i = 0; i < n_out && i < 16; ++i) out[i] = vals[i];
  return ED_OK;
}

int ed_plan_permutation(const ed_plan* P, int64_t* perm_host) {
  if (!P || !perm_host) return fail(ED_ERR_ARG, "ed_plan_permutation: NULL argument");
  std::memcpy(perm_host, P->host.perm.data(), sizeof(int64_t) * P->n);
  return ED_OK;
}

int ed_fmm_matvec_f32(ed_plan* P, const float* sigma, const float* selfi, float* out,
                      ed_stream stream) {
  if (!P || !sigma || !out) return fail(ED_ERR_ARG, "ed_fmm_matvec_f32: NULL argument");
  return fmm_launch(P, sigma, 0, 0, selfi, out, 0, 0, (hipStream_t)stream);
}

int ed_fmm_matvec_f64(ed_plan* P, const double* sigma, const double* selfi, double* out,
                      ed_stream stream) {
  if (!P || !sigma || !out) return fail(ED_ERR_ARG, "ed_fmm_matvec_f64: NULL argument");
  return fmm_launch(P, sigma, 1, 0, selfi, out, 1, 0, (hipStream_t)stream);
}

int ed_fmm_matvec_multi_f64(ed_plan* P, const double* sigma, const double* selfi, double* out,
                            int32_t nrhs, ed_stream stream) {
  if (!P || !sigma || !out) return fail(ED_ERR_ARG, "ed_fmm_matvec_multi_f64: NULL argument");
  if (nrhs < 1 || nrhs > 8) return fail(ED_ERR_ARG, "ed_fmm_matvec_multi_f64: nrhs must be 1..8");
  if (nrhs == 1) return fmm_launch(P, sigma, 1, 0, selfi, out, 1, 0, (hipStream_t)stream);
  ED_TRY(multi_reserve(P, nrhs));
  return fmm_launch(P, sigma, 1, 0, selfi, out, 1, 0, (hipStream_t)stream, nrhs);
}

int ed_plan_attach_near(ed_plan* P, const int64_t* pairs, const double* w, int64_t np) {
  if (!P) return fail(ED_ERR_ARG, "ed_plan_attach_near: NULL plan");
  if (np < 0 || (np > 0 && (!pairs || !w))) return fail(ED_ERR_ARG, "ed_plan_attach_near: bad pairs");
  const int64_t n = P->n;
  std::vector<int64_t> inv(n);
  for (int64_t i = 0; i < n; ++i) inv[P->host.perm[i]] = i;
  struct E { int64_t row, col_orig; int32_t col; double w; };
  std::vector<E> es;
  es.reserve(np);
  for (int64_t q = 0; q < np; ++q) {
    const int64_t i = pairs[2 * q], j = pairs[2 * q + 1];
    if (i < 0 || i >= n || j < 0 || j >= n) continue;  // core/bem.py:709-715 drops invalid
    es.push_back({inv[i], j, (int32_t)inv[j], w[q]});
  }
  std::stable_sort(es.begin(), es.end(), [](const E& a, const E& b) { return a.row < b.row; });
  gmres_invalidate_graph(P);  // a captured Arnoldi graph holds the old near arrays
  sai_free(P);    // the sparse approximate inverse was fitted to the old operator
  std::vector<int32_t> ptr(n + 1, 0), col(es.size());
  std::vector<float> ww(es.size());
  std::vector<double> ww64(es.size());
  for (size_t q = 0; q < es.size(); ++q) {
    ptr[es[q].row + 1]++;
    col[q] = es[q].col;
    ww[q] = (float)es[q].w;
    ww64[q] = es[q].w;
  }
  for (int64_t i = 0; i < n; ++i) ptr[i + 1] += ptr[i];
  (void)hipFree(P->near_ptr);
  (void)hipFree(P->near_col);
  (void)hipFree(P->near_w);
  (void)hipFree(P->near_w64);
  P->near_ptr = nullptr; P->near_col = nullptr; P->near_w = nullptr; P->near_w64 = nullptr;
  ED_TRY(upload(&P->near_ptr, ptr));
  ED_TRY(upload(&P->near_col, col));
  ED_TRY(upload(&P->near_w, ww));
  ED_TRY(upload(&P->near_w64, ww64));
  P->n_near = (int64_t)es.size();
  return ED_OK;
}

int ed_host_tree_build(const double* C, int64_t n, int32_t leaf_size, int32_t max_level,
                       int32_t ws, int64_t* counts, int64_t* perm, int32_t* node_level,
                       int64_t* node_begin, int64_t* node_end, int64_t* node_parent,
                       int64_t* v_pairs, int64_t* u_pairs) {
  if (!C || !counts) return fail(ED_ERR_ARG, "ed_host_tree_build: NULL argument");
  HostTree t;
  ED_TRY(build_host_tree(C, n, leaf_size, max_level, ws, t));
  const int64_t M = (int64_t)t.level.size();
  counts[0] = M;
  counts[1] = t.n_leaves();
  counts[2] = (int64_t)t.v_src.size();
  counts[3] = (int64_t)t.u_src.size();
  if (perm) std::memcpy(perm, t.perm.data(), sizeof(int64_t) * n);
  if (node_level) for (int64_t i = 0; i < M; ++i) node_level[i] = t.level[i];
  if (node_begin) std::memcpy(node_begin, t.begin.data(), sizeof(int64_t) * M);
  if (node_end) std::memcpy(node_end, t.end.data(), sizeof(int64_t) * M);
  if (node_parent) std::memcpy(node_parent, t.parent.data(), sizeof(int64_t) * M);
  if (v_pairs)
    for (int64_t tt = 0; tt < M; ++tt)
      for (int64_t q = t.v_ptr[tt]; q < t.v_ptr[tt + 1]; ++q) {
        v_pairs[2 * q] = tt;
        v_pairs[2 * q + 1] = t.v_src[q];
      }
  if (u_pairs)
    for (int64_t tt = 0; tt < M; ++tt)
      for (int64_t q = t.u_ptr[tt]; q < t.u_ptr[tt + 1]; ++q) {
        u_pairs[2 * q] = tt;
        u_pairs[2 * q + 1] = t.u_src[q];
      }
  return ED_OK;
}

int ed_host_operator(int32_t order, int32_t kind, int32_t ox, int32_t oy, int32_t oz,
                     int32_t octant, double* out) {
  if (!out || order < 1 || order > 16) return fail(ED_ERR_ARG, "ed_host_operator: bad argument");
  if (kind == 0) {
    if (std::max(std::abs(ox), std::max(std::abs(oy), std::abs(oz))) < 2)
      return fail(ED_ERR_ARG, "ed_host_operator: M2L offset must be well separated");
    m2l_matrix(order, ox, oy, oz, out);
  } else if (kind == 1 || kind == 2) {
    if (octant < 0 || octant > 7) return fail(ED_ERR_ARG, "ed_host_operator: octant");
    if (kind == 1) m2m_matrix(order, octant, out);
    else l2l_matrix(order, octant, out);
  } else {
    return fail(ED_ERR_ARG, "ed_host_operator: kind must be 0, 1 or 2");
  }
  return ED_OK;
}

int ed_host_morton_keys_f64(const double* pts, int64_t n, int32_t level, const double* bc,
                            const double* bh, int64_t* keys) {
  if ((!pts && n > 0) || !keys) return fail(ED_ERR_ARG, "ed_host_morton_keys_f64: NULL");
  if (level < 1 || level > 21) return fail(ED_ERR_ARG, "level must be in [1, 21]");
  if ((bc == nullptr) != (bh == nullptr)) return fail(ED_ERR_ARG, "bbox center/half mismatch");
  morton_keys_host(pts, n, level, bc, bh, keys);
  return ED_OK;
}

}  // extern "C"
