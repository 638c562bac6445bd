// C-ABI of the MI355X micromagnetics library: context set-up (orientation, element
// ordering, block partition, device uploads) and the operator / solver entry points.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <numeric>
#include <atomic>
#include <thread>

#include "mm_device.h"
#include "mm_solver.h"
#include "../../include/mumag_hip.h"

namespace mm {

static thread_local std::string g_err;
void set_error(const std::string& msg) { g_err = msg; }

// Scratch of the allocation-free energy entry point (mm_energy_grad_noalloc): the per-block
// (sum, compensation) partials of the energy, 2 doubles per magnetic element block, in a
// file-scope device array of 512 KiB that every GPU the library runs on holds until the process
// ends. Host code never uses the array's name as an address: the device address is resolved
// once per device with hipGetSymbolAddress (with that device current, when the first context
// on it is created) and kept in g_scratch_dev.
constexpr long kScratchDoubles = 1L << 16;
static __device__ __align__(256) double scratch[kScratchDoubles];
static std::mutex g_scratch_mu;
static std::vector<double*> g_scratch_dev;  // device id -> address of that device's scratch

static int resolve_scratch(int device) {
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if ((int)g_scratch_dev.size() <= device) g_scratch_dev.resize(device + 1, nullptr);
  if (g_scratch_dev[device]) return 0;
  double* ptr = nullptr;
  MM_HIP(hipGetSymbolAddress((void**)&ptr, HIP_SYMBOL(scratch)));
  MM_CHECK(ptr != nullptr, "scratch array has no device address");
  g_scratch_dev[device] = ptr;
  return 0;
}

static double* scratch_on(int device) {
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  return device < (int)g_scratch_dev.size() ? g_scratch_dev[device] : nullptr;
}

static std::mutex g_mu;
static std::vector<Ctx*> g_ctx;  // index + 1 is the public handle

static Ctx* get_ctx(int64_t h) {
  std::lock_guard<std::mutex> lk(g_mu);
  if (h <= 0 || h > (int64_t)g_ctx.size() || g_ctx[h - 1] == nullptr) return nullptr;
  return g_ctx[h - 1];
}

// Every fill or copy into a new buffer runs on the context's stream and has completed before the
// buffer is handed out. The stream is created non-blocking, so it does not wait for null-stream
// work: a null-stream hipMemset (asynchronous to the host) issued here could land after the
// stream's first kernel had written the buffer and wipe its results.
template <typename T>
static int dalloc_raw(Ctx& c, T** p, size_t n) {
  if (n == 0) n = 1;
  void* q = nullptr;
  MM_HIP(hipMalloc(&q, n * sizeof(T)));
  c.allocs.push_back(q);
  *p = (T*)q;
  return 0;
}

template <typename T>
static int dalloc(Ctx& c, T** p, size_t n) {
  MM_TRY(dalloc_raw(c, p, n));
  // zeroed once: a segmented solve leaves the entries of segments it does not iterate
  // untouched, and they must hold finite values (the coarsest level multiplies whole vectors)
  MM_HIP(hipMemsetAsync(*p, 0, std::max<size_t>(n, 1) * sizeof(T), c.stream));
  MM_HIP(hipStreamSynchronize(c.stream));
  return 0;
}

template <typename T>
static int upload(Ctx& c, T** p, const T* h, size_t n) {
  if (n == 0) return dalloc(c, p, n);
  MM_TRY(dalloc_raw(c, p, n));
  MM_HIP(hipMemcpyAsync(*p, h, n * sizeof(T), hipMemcpyHostToDevice, c.stream));
  MM_HIP(hipStreamSynchronize(c.stream));
  return 0;
}

// Host-side description of a partition before upload.
struct HostPart {
  std::vector<int> elem_ids;  // original element id of each partition slot
  std::vector<int> eptr, nptr;
  std::vector<uint32_t> nodes, lconn;
  std::vector<uint8_t> mat;
  int max_nodes = 0;
};

// Greedy cut of the ordered element list `elems` into blocks of consecutive elements that
// touch at most kMaxBlockNodes distinct nodes; node lists are sorted; flags mark the first
// block touching a node (owner), nodes touched by one block only (private) and Dirichlet nodes.
static void build_partition(int N, const std::vector<int>& conn, const std::vector<int>& mat0,
                            const std::vector<int>& elems, const std::vector<char>& is_bnd, HostPart& hp,
                            int node_cap, const std::vector<int>& seg_of) {
  int blk_seg = -1;
  std::vector<int> stamp(N, -1), local(N, 0), touches(N, 0);
  std::vector<char> owned(N, 0);
  hp.eptr.assign(1, 0);
  hp.nptr.assign(1, 0);
  hp.nodes.clear();
  hp.lconn.clear();
  hp.mat.clear();
  hp.elem_ids.clear();
  hp.elem_ids.reserve(elems.size());
  hp.lconn.reserve(elems.size());
  hp.mat.reserve(elems.size());
  std::vector<int> blk_nodes;
  int blk = 0;
  size_t blk_e0 = 0;
  auto close_block = [&]() {
    // sort node list, remap local indices of the block's elements
    const int nb = (int)blk_nodes.size();
    std::vector<int> order(nb);
    std::iota(order.begin(), order.end(), 0);
    std::sort(order.begin(), order.end(), [&](int a, int b) { return blk_nodes[a] < blk_nodes[b]; });
    std::vector<int> remap(nb);
    for (int i = 0; i < nb; ++i) remap[order[i]] = i;
    for (size_t e = blk_e0; e < hp.lconn.size(); ++e) {
      uint32_t lc = hp.lconn[e], out = 0;
      for (int a = 0; a < 4; ++a) out |= (uint32_t)remap[(lc >> (8 * a)) & 255u] << (8 * a);
      hp.lconn[e] = out;
    }
    // Interleave the block's elements (a 64 x s transpose) so that the 64 elements a wave
    // processes together lie s positions apart in the (min node)-sorted order: they touch
    // different nodes and their LDS atomics rarely collide.
    {
      const int n = (int)(hp.lconn.size() - blk_e0);
      const int st = (n + 63) / 64;
      std::vector<int> perm;
      perm.reserve(n);
      for (int cc = 0; cc < st; ++cc)
        for (int lane = 0; lane < 64; ++lane) {
          const int old = lane * st + cc;
          if (old < n) perm.push_back(old);
        }
      std::vector<uint32_t> lc2(n);
      std::vector<uint8_t> mt2(n);
      std::vector<int> id2(n);
      for (int j = 0; j < n; ++j) {
        lc2[j] = hp.lconn[blk_e0 + perm[j]];
        mt2[j] = hp.mat[blk_e0 + perm[j]];
        id2[j] = hp.elem_ids[blk_e0 + perm[j]];
      }
      for (int j = 0; j < n; ++j) {
        hp.lconn[blk_e0 + j] = lc2[j];
        hp.mat[blk_e0 + j] = mt2[j];
        hp.elem_ids[blk_e0 + j] = id2[j];
      }
    }
    for (int i = 0; i < nb; ++i) {
      const int v = blk_nodes[order[i]];
      uint32_t f = (uint32_t)v;
      if (!owned[v]) {
        owned[v] = 1;
        f |= kOwner;
      }
      if (is_bnd[v]) f |= kBoundary;
      touches[v]++;
      hp.nodes.push_back(f);
    }
    hp.max_nodes = std::max(hp.max_nodes, nb);
    hp.eptr.push_back((int)hp.lconn.size());
    hp.nptr.push_back((int)hp.nodes.size());
    blk_nodes.clear();
    blk_e0 = hp.lconn.size();
    ++blk;
  };
  for (size_t k = 0; k < elems.size(); ++k) {
    const int e = elems[k];
    int fresh = 0;
    for (int a = 0; a < 4; ++a) {
      const int v = conn[4 * e + a];
      bool dup = false;
      for (int b2 = 0; b2 < a; ++b2) dup |= conn[4 * e + b2] == v;
      if (!dup && stamp[v] != blk) ++fresh;
    }
    const int sg = seg_of[conn[4 * e]];
    if (!blk_nodes.empty() && ((int)blk_nodes.size() + fresh > node_cap ||
                               (int)(hp.lconn.size() - blk_e0) >= kMaxBlockElems || sg != blk_seg))
      close_block();
    blk_seg = sg;
    uint32_t lc = 0;
    for (int a = 0; a < 4; ++a) {
      const int v = conn[4 * e + a];
      if (stamp[v] != blk) {
        stamp[v] = blk;
        local[v] = (int)blk_nodes.size();
        blk_nodes.push_back(v);
      }
      lc |= (uint32_t)local[v] << (8 * a);
    }
    hp.lconn.push_back(lc);
    hp.mat.push_back((uint8_t)mat0[e]);
    hp.elem_ids.push_back(e);
  }
  if (!blk_nodes.empty()) close_block();
  for (auto& f : hp.nodes)
    if (touches[f & kIdMask] == 1) f |= kPrivate;
}

// ------------------------------------------------------------------ edge-star clusters
// The Poisson apply reads every corner's coordinates and value from LDS; per tet that is
// 16 doubles plus 4 LDS atomics, and LDS bandwidth is what bounds the per-tet kernel.
// Tets that share an edge share its two nodes, and neighbours around the edge share a
// third, so the apply processes clusters: up to 6 tets on 8 node slots [E0, E1, R0..R5],
// tet j having corners (E0, E1, Rj, R(j+1) mod 6). One thread loads the 8 slots once,
// keeps their sums in registers and issues 8 LDS atomics for the whole cluster.
// Tets are grouped by their longest edge (ties: the smaller node pair); on a six-tet hex
// split that edge is the hex diagonal shared by all six tets, so a cluster is a whole hex.
// Around the edge, tets that share a face are chained so consecutive tets share a slot.
struct HostCluster {
  int node[8];      // global node ids, -1 for an unused slot
  uint32_t tets;    // bit j: tet j present
};

static void build_clusters(const double* knt, int E, const std::vector<int>& conn, std::vector<HostCluster>& out) {
  std::vector<std::pair<uint64_t, int>> key(E);
  for (int e = 0; e < E; ++e) {
    const int* v = &conn[4 * (size_t)e];
    double best = -1.0;
    uint64_t bk = 0;
    for (int a = 0; a < 4; ++a)
      for (int b = a + 1; b < 4; ++b) {
        double l2 = 0.0;
        for (int k = 0; k < 3; ++k) {
          const double d = knt[3 * (size_t)v[a] + k] - knt[3 * (size_t)v[b] + k];
          l2 += d * d;
        }
        const uint64_t lo = (uint64_t)std::min(v[a], v[b]), hi = (uint64_t)std::max(v[a], v[b]);
        const uint64_t k2 = (lo << 32) | hi;
        if (l2 > best || (l2 == best && k2 < bk)) {
          best = l2;
          bk = k2;
        }
      }
    key[e] = {bk, e};
  }
  std::sort(key.begin(), key.end());
  out.clear();
  out.reserve(E / 4 + 16);
  std::vector<int> T, opp0, opp1;
  std::vector<char> used;
  std::vector<std::vector<int>> chains;
  for (size_t g0 = 0; g0 < key.size();) {
    size_t g1 = g0;
    while (g1 < key.size() && key[g1].first == key[g0].first) ++g1;
    const int a = (int)(key[g0].first >> 32), b = (int)(key[g0].first & 0xffffffffu);
    const int n = (int)(g1 - g0);
    T.resize(n);
    opp0.resize(n);
    opp1.resize(n);
    for (int i = 0; i < n; ++i) {
      const int e = key[g0 + i].second;
      T[i] = e;
      int o[2], k = 0;
      for (int q = 0; q < 4; ++q) {
        const int v = conn[4 * (size_t)e + q];
        if (v != a && v != b && k < 2) o[k++] = v;
      }
      opp0[i] = o[0];
      opp1[i] = o[1];
    }
    // chains of face-sharing tets around the edge: ring-node sequences (tets = nodes - 1)
    auto degree = [&](int v) {
      int d = 0;
      for (int i = 0; i < n; ++i) d += (opp0[i] == v) + (opp1[i] == v);
      return d;
    };
    used.assign(n, 0);
    chains.clear();
    bool closed6 = false;
    for (int left = n; left > 0;) {
      int start = -1, snode = -1;
      for (int i = 0; i < n && start < 0; ++i)
        if (!used[i]) {
          if (degree(opp0[i]) == 1) start = i, snode = opp0[i];
          else if (degree(opp1[i]) == 1) start = i, snode = opp1[i];
        }
      if (start < 0)
        for (int i = 0; i < n && start < 0; ++i)
          if (!used[i]) start = i, snode = opp0[i];
      std::vector<int> ch{snode};
      int t = start, cur = snode;
      while (t >= 0) {
        used[t] = 1;
        --left;
        cur = (opp0[t] == cur) ? opp1[t] : opp0[t];
        ch.push_back(cur);
        t = -1;
        for (int i = 0; i < n; ++i)
          if (!used[i] && (opp0[i] == cur || opp1[i] == cur)) {
            t = i;
            break;
          }
      }
      if (ch.size() == 7 && ch.front() == ch.back() && n == 6) closed6 = true;
      chains.push_back(std::move(ch));
    }
    if (closed6) {
      HostCluster hc;
      hc.node[0] = a;
      hc.node[1] = b;
      for (int s2 = 0; s2 < 6; ++s2) hc.node[2 + s2] = chains[0][s2];
      hc.tets = 63u;
      out.push_back(hc);
    } else {
      // pieces of at most 5 tets (6 ring nodes), packed onto the 6 ring slots
      HostCluster hc;
      int used_slots = 0;
      auto flush = [&]() {
        if (used_slots == 0) return;
        hc.node[0] = a;
        hc.node[1] = b;
        for (int s2 = used_slots; s2 < 6; ++s2) hc.node[2 + s2] = -1;
        out.push_back(hc);
        used_slots = 0;
        hc.tets = 0;
      };
      hc.tets = 0;
      for (const auto& ch : chains)
        for (size_t p0 = 0; p0 + 1 < ch.size(); p0 += 5) {
          const size_t p1 = std::min(ch.size(), p0 + 6);  // nodes [p0, p1)
          const int len = (int)(p1 - p0);
          if (used_slots + len > 6) flush();
          for (int q = 0; q < len; ++q) hc.node[2 + used_slots + q] = ch[p0 + q];
          for (int q = 0; q + 1 < len; ++q) hc.tets |= 1u << (used_slots + q);
          used_slots += len;
        }
      flush();
    }
    g0 = g1;
  }
}

// Greedy cut of the clusters (ordered by smallest, then largest node id) into blocks touching
// at most node_cap distinct nodes, with the flags of build_partition. Per cluster: w0/w1 hold
// the block-local indices of slots 0-3 / 4-7 (low 8 bits each), wm bits 0-5 the tets and bits
// 8 + 3q .. 10 + 3q the high bits of slot q's index; the kernels derive the used slots from the
// tets (slots 0, 1 and ring slot 2 + k of tets k and k - 1 mod 6), which upload_cluster_part
// checks against the slots that hold nodes.
struct HostCPart {
  std::vector<int> cptr, nptr;
  std::vector<uint32_t> nodes, w0, w1, wm;
  int max_nodes = 0;
};

static void build_cluster_partition(int N, std::vector<HostCluster>& cl, const std::vector<char>& is_bnd, HostCPart& hp,
                                    int node_cap, const std::vector<int>& seg_of) {
  int blk_seg = -1;
  std::vector<std::pair<uint64_t, int>> key(cl.size());
  for (size_t i = 0; i < cl.size(); ++i) {
    int lo = INT32_MAX, hi = -1;
    for (int q = 0; q < 8; ++q)
      if (cl[i].node[q] >= 0) {
        lo = std::min(lo, cl[i].node[q]);
        hi = std::max(hi, cl[i].node[q]);
      }
    key[i] = {((uint64_t)lo << 32) | (uint64_t)hi, (int)i};
  }
  std::stable_sort(key.begin(), key.end());
  std::vector<int> stamp(N, -1), local(N, 0), touches(N, 0);
  std::vector<char> owned(N, 0);
  std::vector<int> blk_nodes;
  hp.cptr.assign(1, 0);
  hp.nptr.assign(1, 0);
  int blk = 0;
  std::vector<int> members;
  auto close_block = [&]() {
    const int nb = (int)blk_nodes.size();
    std::vector<int> order(nb);
    std::iota(order.begin(), order.end(), 0);
    std::sort(order.begin(), order.end(), [&](int x, int y) { return blk_nodes[x] < blk_nodes[y]; });
    std::vector<int> remap(nb);
    for (int i = 0; i < nb; ++i) remap[order[i]] = i;
    for (int ci : members) {
      const HostCluster& h = cl[ci];
      uint32_t w[2] = {0u, 0u}, m = h.tets;
      for (int q = 0; q < 8; ++q)
        if (h.node[q] >= 0) {
          const uint32_t li = (uint32_t)remap[local[h.node[q]]];
          w[q >> 2] |= (li & 255u) << (8 * (q & 3));
          m |= (li >> 8) << (8 + 3 * q);
        }
      hp.w0.push_back(w[0]);
      hp.w1.push_back(w[1]);
      hp.wm.push_back(m);
    }
    for (int i = 0; i < nb; ++i) {
      const int v = blk_nodes[order[i]];
      uint32_t f = (uint32_t)v;
      if (!owned[v]) {
        owned[v] = 1;
        f |= kOwner;
      }
      if (is_bnd[v]) f |= kBoundary;
      touches[v]++;
      hp.nodes.push_back(f);
    }
    hp.max_nodes = std::max(hp.max_nodes, nb);
    hp.cptr.push_back((int)hp.w0.size());
    hp.nptr.push_back((int)hp.nodes.size());
    blk_nodes.clear();
    members.clear();
    ++blk;
  };
  for (size_t k = 0; k < key.size(); ++k) {
    const HostCluster& h = cl[key[k].second];
    int fresh = 0;
    for (int q = 0; q < 8; ++q) {
      const int v = h.node[q];
      if (v < 0) continue;
      bool dup = false;
      for (int q2 = 0; q2 < q; ++q2) dup |= h.node[q2] == v;
      if (!dup && stamp[v] != blk) ++fresh;
    }
    const int sg = seg_of[h.node[0]];
    if (!blk_nodes.empty() &&
        ((int)blk_nodes.size() + fresh > node_cap || (int)members.size() >= kMaxBlockElems || sg != blk_seg))
      close_block();
    blk_seg = sg;
    for (int q = 0; q < 8; ++q) {
      const int v = h.node[q];
      if (v >= 0 && stamp[v] != blk) {
        stamp[v] = blk;
        local[v] = (int)blk_nodes.size();
        blk_nodes.push_back(v);
      }
    }
    members.push_back(key[k].second);
  }
  if (!blk_nodes.empty()) close_block();
  for (auto& f : hp.nodes)
    if (touches[f & kIdMask] == 1) f |= kPrivate;
}

// Cut the context's clusters into blocks of at most `cap` distinct nodes and upload them as
// part_cl with their entry records, freeing a previous cluster partition. hc receives the host
// partition; knt_host: the node coordinates, or null to read them back from the device.
static int upload_cluster_part(Ctx& c, int cap, const std::vector<int>& seg_of, HostCPart& hc,
                               const double* knt_host) {
  std::vector<HostCluster> cl(c.h_cl_tets.size());
  for (size_t i = 0; i < cl.size(); ++i) {
    for (int q = 0; q < 8; ++q) cl[i].node[q] = c.h_cl_nodes[8 * i + q];
    cl[i].tets = c.h_cl_tets[i];
  }
  for (const HostCluster& h : cl) {
    uint32_t held = 0u;
    for (int q = 0; q < 8; ++q)
      if (h.node[q] >= 0) held |= 1u << q;
    const uint32_t t = h.tets & 63u, ring = (t | (t << 1) | (t >> 5)) & 63u;
    MM_CHECK(t != 0u && held == (3u | (ring << 2)), "cluster slots do not follow from its tets");
  }
  build_cluster_partition(c.N, cl, c.h_is_bnd, hc, cap, seg_of);
  // entry records of the FP32 kernels in block order: coordinates relative to the block's first
  // node (subtracted in FP64, then rounded, as the gathering staging does) with id | flags
  std::vector<double> hk;
  if (!knt_host) {
    hk.resize(3 * (size_t)c.N);
    MM_HIP(hipMemcpyAsync(hk.data(), c.knt, hk.size() * sizeof(double), hipMemcpyDeviceToHost, c.stream));
    MM_HIP(hipStreamSynchronize(c.stream));
    knt_host = hk.data();
  }
  const double* kh = knt_host;
  const size_t n_ent = hc.nodes.size();
  std::vector<float4> r32(n_ent);
  for (size_t b = 0; b + 1 < hc.nptr.size(); ++b) {
    if (hc.nptr[b] == hc.nptr[b + 1]) continue;
    const size_t id0 = hc.nodes[hc.nptr[b]] & kIdMask;
    const double ox = kh[3 * id0], oy = kh[3 * id0 + 1], oz = kh[3 * id0 + 2];
    for (int n = hc.nptr[b]; n < hc.nptr[b + 1]; ++n) {
      const uint32_t f = hc.nodes[n];
      const size_t id = f & kIdMask;
      float4 g;
      g.x = (float)(kh[3 * id] - ox);
      g.y = (float)(kh[3 * id + 1] - oy);
      g.z = (float)(kh[3 * id + 2] - oz);
      std::memcpy(&g.w, &f, sizeof(float));
      r32[n] = g;
    }
  }
  Part& q = c.part_cl;
  for (const void* old : {(const void*)q.eptr, (const void*)q.nptr, (const void*)q.nodes, (const void*)q.lconn,
                          (const void*)q.lconn2, (const void*)q.cmask, (const void*)q.rec32})
    if (old) {
      auto it = std::find(c.allocs.begin(), c.allocs.end(), old);
      if (it != c.allocs.end()) {
        (void)hipFree(*it);
        c.allocs.erase(it);
      }
    }
  q = Part();
  q.nblk = (int)hc.cptr.size() - 1;
  q.E = (int)hc.w0.size();
  int *cp, *np;
  uint32_t *nodes, *w0, *w1, *wm;
  MM_TRY(upload(c, &cp, hc.cptr.data(), hc.cptr.size()));
  MM_TRY(upload(c, &np, hc.nptr.data(), hc.nptr.size()));
  MM_TRY(upload(c, &nodes, hc.nodes.data(), hc.nodes.size()));
  MM_TRY(upload(c, &w0, hc.w0.data(), hc.w0.size()));
  MM_TRY(upload(c, &w1, hc.w1.data(), hc.w1.size()));
  MM_TRY(upload(c, &wm, hc.wm.data(), hc.wm.size()));
  float4* d32;
  MM_TRY(upload(c, &d32, r32.data(), r32.size()));
  q.eptr = cp;
  q.nptr = np;
  q.nodes = nodes;
  q.lconn = w0;
  q.lconn2 = w1;
  q.cmask = wm;
  q.rec32 = d32;
  c.nodes_cl = (long long)hc.nodes.size();
  c.cl_cap = cap;
  return 0;
}

// Greedy colouring of the cluster blocks for the colour-sequenced scatter: blocks that share a
// node get different colours (each block takes the smallest colour none of its already coloured
// neighbours holds, in block order, so a colour's blocks stay in spatial order). The blocks are
// uploaded colour by colour; the colours run as consecutive launches on the context's stream.
int build_block_colours(Ctx& c) {
  const Part& p = c.part_cl;
  const int nb = p.nblk;
  std::vector<int> nptr(nb + 1);
  std::vector<uint32_t> nodes((size_t)c.nodes_cl);
  MM_HIP(hipStreamSynchronize(c.stream));
  MM_HIP(hipMemcpyAsync(nptr.data(), p.nptr, nptr.size() * sizeof(int), hipMemcpyDeviceToHost, c.stream));
  if (!nodes.empty())
    MM_HIP(hipMemcpyAsync(nodes.data(), p.nodes, nodes.size() * sizeof(uint32_t), hipMemcpyDeviceToHost, c.stream));
  MM_HIP(hipStreamSynchronize(c.stream));
  // blocks touching each shared node, in block order (CSR over the nodes)
  std::vector<long long> off((size_t)c.N + 1, 0);
  for (int b = 0; b < nb; ++b)
    for (int e = nptr[b]; e < nptr[b + 1]; ++e)
      if (!(nodes[e] & kPrivate)) off[(nodes[e] & kIdMask) + 1]++;
  for (int v = 0; v < c.N; ++v) off[v + 1] += off[v];
  std::vector<int> blk((size_t)off[c.N]);
  std::vector<long long> fill(off.begin(), off.end() - 1);
  for (int b = 0; b < nb; ++b)
    for (int e = nptr[b]; e < nptr[b + 1]; ++e)
      if (!(nodes[e] & kPrivate)) blk[fill[nodes[e] & kIdMask]++] = b;
  std::vector<int> colour(nb, -1), stamp;
  int ncol = 0;
  for (int b = 0; b < nb; ++b) {
    for (int e = nptr[b]; e < nptr[b + 1]; ++e) {
      if (nodes[e] & kPrivate) continue;
      const uint32_t v = nodes[e] & kIdMask;
      for (long long k = off[v]; k < off[v + 1]; ++k) {
        const int cb = colour[blk[k]];
        if (cb >= 0) stamp[cb] = b;  // a neighbour coloured earlier
      }
    }
    int col = 0;
    while (col < ncol && stamp[col] == b) ++col;
    if (col == ncol) {
      stamp.push_back(-1);
      ++ncol;
    }
    colour[b] = col;
  }
  std::vector<int> coff(ncol + 1, 0), order(nb);
  for (int b = 0; b < nb; ++b) coff[colour[b] + 1]++;
  for (int k = 0; k < ncol; ++k) coff[k + 1] += coff[k];
  std::vector<int> at(coff.begin(), coff.end() - 1);
  for (int b = 0; b < nb; ++b) order[at[colour[b]]++] = b;
  if (c.d_colour_blk) {
    auto it = std::find(c.allocs.begin(), c.allocs.end(), (void*)c.d_colour_blk);
    if (it != c.allocs.end()) {
      (void)hipFree(*it);
      c.allocs.erase(it);
    }
    c.d_colour_blk = nullptr;
  }
  MM_TRY(upload(c, &c.d_colour_blk, order.data(), order.size()));
  c.h_colour_off = coff;
  return 0;
}

static int upload_part(Ctx& c, const HostPart& hp, int E, Part& p) {
  p.nblk = (int)hp.eptr.size() - 1;
  p.E = E;
  p.Es = ((E + 63) / 64) * 64;
  int *ep, *np;
  uint32_t *nodes, *lconn;
  uint8_t* mat;
  MM_TRY(upload(c, &ep, hp.eptr.data(), hp.eptr.size()));
  MM_TRY(upload(c, &np, hp.nptr.data(), hp.nptr.size()));
  MM_TRY(upload(c, &nodes, hp.nodes.data(), hp.nodes.size()));
  MM_TRY(upload(c, &lconn, hp.lconn.data(), hp.lconn.size()));
  MM_TRY(upload(c, &mat, hp.mat.data(), hp.mat.size()));
  p.eptr = ep;
  p.nptr = np;
  p.nodes = nodes;
  p.lconn = lconn;
  p.mat = mat;
  return 0;
}

static int create(int N, int E, const double* knt, const int32_t* conn_in, const int32_t* mat_id, int G,
                  const double* A_red, const double* K1_red, const double* Js_red, const double* k_easy,
                  const double* bnd_mask, double reg, int flags, Ctx& c, int nseg = 0,
                  const int32_t* seg_node_off = nullptr) {
  MM_CHECK(N > 0 && E > 0, "N and E must be positive");
  MM_CHECK(N < (int)kIdMask, "too many nodes");
  MM_CHECK(G >= 1 && G <= kMaxMat, "material count must be in [1, 255]");
  MM_CHECK(knt && conn_in && mat_id && A_red && K1_red && Js_red && k_easy, "null input array");
  c.N = N;
  c.E = E;
  c.G = G;
  c.reg = reg;
  c.geom_mode = flags & 3;
  c.node_cap = (flags >> 8) & 511 ? (flags >> 8) & 511 : kMaxBlockNodes;
  MM_CHECK(c.geom_mode <= 2, "invalid geometry mode");
  MM_HIP(hipGetDevice(&c.device));
  MM_TRY(resolve_scratch(c.device));
  MM_HIP(hipDeviceGetAttribute(&c.num_cu, hipDeviceAttributeMultiprocessorCount, c.device));
  MM_HIP(hipStreamCreateWithFlags(&c.stream, hipStreamNonBlocking));
  MM_HIP(hipEventCreateWithFlags(&c.ev, hipEventDisableTiming));

  // segments: contiguous node ranges of independent systems (one range when none are given)
  std::vector<int> seg_of(N, 0);
  c.h_seg_node_off.assign({0, N});
  if (nseg > 0) {
    MM_CHECK(seg_node_off && seg_node_off[0] == 0 && seg_node_off[nseg] == N, "segment offsets must run from 0 to N");
    c.h_seg_node_off.assign(seg_node_off, seg_node_off + nseg + 1);
    for (int g = 0; g < nseg; ++g) {
      MM_CHECK(seg_node_off[g + 1] > seg_node_off[g], "segments must be non-empty and increasing");
      for (int v = seg_node_off[g]; v < seg_node_off[g + 1]; ++v) seg_of[v] = g;
    }
    for (int e = 0; e < E; ++e) {
      for (int q = 0; q < 4; ++q)
        MM_CHECK(conn_in[4 * (size_t)e + q] >= 0 && conn_in[4 * (size_t)e + q] < N,
                 "connectivity index out of range");
      for (int q = 1; q < 4; ++q)
        MM_CHECK(seg_of[conn_in[4 * (size_t)e + q]] == seg_of[conn_in[4 * (size_t)e]],
                 "an element couples two segments");
    }
  }
  c.nseg = (int)c.h_seg_node_off.size() - 1;

  // orientation fix (src/loop.py:70-85): swap corners 1 and 2 where det J < 0
  std::vector<int> conn(conn_in, conn_in + 4 * (size_t)E);
  std::vector<int> mat0(E);
  for (int e = 0; e < E; ++e) {
    for (int a = 0; a < 4; ++a)
      MM_CHECK(conn[4 * e + a] >= 0 && conn[4 * e + a] < N, "connectivity index out of range");
    MM_CHECK(mat_id[e] >= 1 && mat_id[e] <= G, "material id out of range [1, G]");
    mat0[e] = mat_id[e] - 1;
    const double* x0 = knt + 3 * (size_t)conn[4 * e];
    double cc[3][3];
    for (int b = 0; b < 3; ++b)
      for (int k = 0; k < 3; ++k) cc[b][k] = knt[3 * (size_t)conn[4 * e + b + 1] + k] - x0[k];
    const double det = cc[0][0] * (cc[1][1] * cc[2][2] - cc[1][2] * cc[2][1]) -
                       cc[0][1] * (cc[1][0] * cc[2][2] - cc[1][2] * cc[2][0]) +
                       cc[0][2] * (cc[1][0] * cc[2][1] - cc[1][1] * cc[2][0]);
    // a flat element has no J^-1 (the reference's inversion of J fails there too)
    MM_CHECK(det != 0.0 && std::isfinite(det), "degenerate element (zero or non-finite volume)");
    if (det < 0) std::swap(conn[4 * e + 1], conn[4 * e + 2]);
  }
  // element order: by (smallest node id, largest node id), stable
  std::vector<long long> key(E);
  for (int e = 0; e < E; ++e) {
    int lo = conn[4 * e], hi = conn[4 * e];
    for (int a = 1; a < 4; ++a) {
      lo = std::min(lo, conn[4 * e + a]);
      hi = std::max(hi, conn[4 * e + a]);
    }
    key[e] = ((long long)lo << 31) | hi;
  }
  std::vector<int> order(E);
  std::iota(order.begin(), order.end(), 0);
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return key[a] < key[b]; });

  std::vector<char> is_bnd(N, 0);
  if (bnd_mask)
    for (int i = 0; i < N; ++i) is_bnd[i] = bnd_mask[i] == 0.0;

  // material table
  std::memset(&c.hmt, 0, sizeof(c.hmt));
  for (int g = 0; g < G; ++g) {
    c.hmt.A[g] = A_red[g];
    c.hmt.K1[g] = K1_red[g];
    c.hmt.Js[g] = Js_red[g];
    for (int k = 0; k < 3; ++k) c.hmt.k[g][k] = k_easy[3 * g + k];
  }

  // distinct nodes per element block: flags bits 8..16 (0: 256, the LDS staging capacity)
  int node_cap = (flags >> 8) & 511;
  if (node_cap == 0) node_cap = kMaxBlockNodes;
  MM_CHECK(node_cap >= 16 && node_cap <= kMaxBlockNodes, "block node cap (flags bits 8..16) must be in [16, 256]");
  HostPart hp_all, hp_mag;
  build_partition(N, conn, mat0, order, is_bnd, hp_all, node_cap, seg_of);
  std::vector<int> mag_elems;
  mag_elems.reserve(E);
  for (int k = 0; k < E; ++k) {
    const int g = mat0[order[k]];
    if (A_red[g] != 0.0 || K1_red[g] != 0.0 || Js_red[g] != 0.0) mag_elems.push_back(order[k]);
  }
  const std::vector<char> no_bnd(N, 0);
  build_partition(N, conn, mat0, mag_elems, no_bnd, hp_mag, node_cap, seg_of);

  MM_TRY(upload_part(c, hp_all, E, c.part_all));
  // block offsets per segment; blocks follow the element order (smallest node id first), so a
  // segment's blocks are contiguous
  bool seg_sorted = true;
  auto seg_blocks = [&](const std::vector<int>& nptr, const std::vector<uint32_t>& nodes) {
    std::vector<int> off(c.nseg + 1, 0);
    int prev = 0;
    for (size_t b = 0; b + 1 < nptr.size(); ++b) {
      const int sg = seg_of[nodes[nptr[b]] & kIdMask];
      seg_sorted &= sg >= prev;
      prev = sg;
      off[sg + 1]++;
    }
    for (int g = 0; g < c.nseg; ++g) off[g + 1] += off[g];
    return off;
  };
  c.h_seg_blk_all = seg_blocks(hp_all.nptr, hp_all.nodes);
  {
    std::vector<HostCluster> cl;
    build_clusters(knt, E, conn, cl);
    c.h_cl_nodes.resize(8 * cl.size());
    c.h_cl_tets.resize(cl.size());
    for (size_t i = 0; i < cl.size(); ++i) {
      for (int q = 0; q < 8; ++q) c.h_cl_nodes[8 * i + q] = cl[i].node[q];
      c.h_cl_tets[i] = cl[i].tets;
    }
    c.h_is_bnd = is_bnd;
    HostCPart hc;
    MM_TRY(upload_cluster_part(c, node_cap, seg_of, hc, knt));
    c.h_seg_blk_cl = seg_blocks(hc.nptr, hc.nodes);
  }
  c.h_seg_blk_mag = seg_blocks(hp_mag.nptr, hp_mag.nodes);
  MM_CHECK(seg_sorted, "element blocks are not ordered by segment");
  c.nodes_all = (long long)hp_all.nodes.size();
  MM_TRY(upload_part(c, hp_mag, (int)mag_elems.size(), c.part_mag));
  c.max_block_nodes = std::max(hp_all.max_nodes, hp_mag.max_nodes);
  c.n_mag = (int)mag_elems.size();

  // device geometry in part_all order
  c.h_conn.resize(4 * (size_t)E);
  order = hp_all.elem_ids;
  for (int k = 0; k < E; ++k)
    for (int a = 0; a < 4; ++a) c.h_conn[4 * (size_t)k + a] = conn[4 * (size_t)order[k] + a];
  MM_TRY(upload(c, &c.knt, knt, 3 * (size_t)N));
  int* conn_dev = nullptr;
  MM_TRY(upload(c, &conn_dev, c.h_conn.data(), c.h_conn.size()));
  MatTable* mt = nullptr;
  MM_TRY(upload(c, &mt, &c.hmt, 1));
  const size_t Es = c.part_all.Es;
  double *grad, *Ve;
  float *grad32, *Ve32;
  MM_TRY(dalloc(c, &grad, 9 * Es));
  MM_TRY(dalloc(c, &Ve, Es));
  MM_TRY(dalloc(c, &grad32, 9 * Es));
  MM_TRY(dalloc(c, &Ve32, Es));
  c.geo.knt = c.knt;
  c.geo.grad = grad;
  c.geo.Ve = Ve;
  c.geo.grad32 = grad32;
  c.geo.Ve32 = Ve32;
  c.geo.mt = mt;
  MM_TRY(launch_geometry(c, conn_dev, c.stream));
  MM_HIP(hipGetLastError());
  std::vector<double> hg(9 * Es);
  c.h_Ve.resize(Es);
  MM_HIP(hipMemcpyAsync(hg.data(), grad, 9 * Es * sizeof(double), hipMemcpyDeviceToHost, c.stream));
  MM_HIP(hipMemcpyAsync(c.h_Ve.data(), Ve, Es * sizeof(double), hipMemcpyDeviceToHost, c.stream));
  MM_HIP(hipStreamSynchronize(c.stream));
  c.h_Ve.resize(E);
  c.h_grad.resize(9 * (size_t)E);
  for (int e = 0; e < E; ++e)
    for (int j = 0; j < 9; ++j) c.h_grad[9 * (size_t)e + j] = hg[(size_t)j * Es + e];

  // nodal set-up: V_mag, M_nodal, exchange diagonal (src/loop.py:862-865,929-939;
  // src/energy_kernels.py:492-566; src/hysteresis_loop.py:340-353)
  // (a context of several independent systems keeps one V_mag per segment)
  double V_mag = 0.0;
  c.h_seg_vmag.assign(c.nseg, 0.0);
  for (int k = 0; k < E; ++k)
    if (Js_red[mat0[order[k]]] > 0) {
      V_mag += c.h_Ve[k];
      c.h_seg_vmag[seg_of[c.h_conn[4 * (size_t)k]]] += c.h_Ve[k];
    }
  c.V_mag = V_mag == 0.0 ? 1.0 : V_mag;
  for (double& v : c.h_seg_vmag)
    if (v == 0.0) v = 1.0;
  std::vector<double> vnode(N, c.V_mag);
  if (c.nseg > 1)
    for (int i = 0; i < N; ++i) vnode[i] = c.h_seg_vmag[seg_of[i]];
  c.h_M_nodal.assign(N, 0.0);
  std::vector<double> dex(N, 0.0);
  for (int k = 0; k < E; ++k) {
    const int g = mat0[order[k]];
    const double ve = c.h_Ve[k];
    const double* gr = &c.h_grad[9 * (size_t)k];
    double G4[4][3];
    for (int j = 0; j < 3; ++j) {
      G4[1][j] = gr[j];
      G4[2][j] = gr[3 + j];
      G4[3][j] = gr[6 + j];
      G4[0][j] = -(G4[1][j] + G4[2][j] + G4[3][j]);
    }
    for (int a = 0; a < 4; ++a) {
      const int v = c.h_conn[4 * (size_t)k + a];
      c.h_M_nodal[v] += Js_red[g] * ve * 0.25;
      dex[v] += 2.0 * A_red[g] * ve * (G4[a][0] * G4[a][0] + G4[a][1] * G4[a][1] + G4[a][2] * G4[a][2]);
    }
  }
  c.h_inv_M_rel.resize(N);
  c.h_inv_M_prec.resize(N);
  c.h_exdiag.resize(N);
  for (int i = 0; i < N; ++i) {
    const double M = c.h_M_nodal[i];
    c.h_inv_M_rel[i] = M > 1e-20 ? vnode[i] / M : 0.0;
    const double dd = dex[i] / vnode[i];
    c.h_exdiag[i] = dd;
    c.h_inv_M_prec[i] = dd > 1e-20 ? 1.0 / dd : 1.0;
  }
  MM_TRY(upload(c, &c.M_nodal, c.h_M_nodal.data(), N));
  MM_TRY(upload(c, &c.inv_M_rel, c.h_inv_M_rel.data(), N));
  MM_TRY(upload(c, &c.inv_M_prec, c.h_inv_M_prec.data(), N));
  c.geo.M_nodal = c.M_nodal;
  if (c.nseg > 1) {
    for (double& v : vnode) v = 1.0 / v;
    MM_TRY(upload(c, &c.inv_V_node, vnode.data(), N));
    c.geo.inv_V_node = c.inv_V_node;
    int* d = nullptr;
    MM_TRY(upload(c, &d, c.h_seg_blk_mag.data(), c.h_seg_blk_mag.size()));
    c.d_seg_blk_mag = d;
  }

  // reductions and scalars
  const int max_grid = std::max({kRedBlocks, 8 * ((c.part_all.nblk + 7) / 8), 8 * ((c.part_mag.nblk + 7) / 8)});
  MM_TRY(dalloc(c, &c.partial, 4 * (size_t)max_grid));
  MM_TRY(dalloc(c, &c.dscal, kNumScal));
  MM_HIP(hipHostMalloc((void**)&c.hscal, kNumScal * sizeof(double), hipHostMallocDefault));
  MM_TRY(dalloc(c, &c.part_pq, (size_t)8 * ((std::max(c.part_all.nblk, c.part_cl.nblk) + 7) / 8) + kRedGridMax));
  MM_TRY(dalloc(c, &c.part_rr, 2 * (size_t)kRedGridMax));
  MM_TRY(dalloc(c, &c.part_rz, (size_t)kRedGridMax));
  // r.r partials go straight from the PCG kernels into host-coherent pinned memory: the host
  // reads them after the iteration's event, with no copy on the stream
  MM_HIP(hipHostMalloc((void**)&c.hpart, 2 * kRedGridMax * sizeof(double), hipHostMallocMapped | hipHostMallocCoherent));
  MM_HIP(hipHostGetDevicePointer((void**)&c.hpart_dev, c.hpart, 0));

  // Poisson work vectors
  for (double** v : {&c.pb, &c.pr, &c.pz, &c.pp, &c.pq, &c.pt, &c.px0, &c.work_U, &c.diagA}) MM_TRY(dalloc(c, v, N));
  MM_TRY(dalloc(c, &c.px0f, N));
  MM_TRY(dalloc(c, &c.ptf, N));
  MM_HIP(hipStreamSynchronize(c.stream));
  return 0;
}

static void destroy(Ctx* c) {
  if (!c) return;
  (void)hipDeviceSynchronize();
  for (void* p : c->allocs) (void)hipFree(p);
  if (c->hscal) (void)hipHostFree(c->hscal);
  if (c->hpart) (void)hipHostFree(c->hpart);
  if (c->hseg) (void)hipHostFree(c->hseg);
  if (c->hguess) (void)hipHostFree(c->hguess);
  if (c->hmpart) (void)hipHostFree(c->hmpart);
  if (c->seg_stage) (void)hipHostFree(c->seg_stage);
  for (hipEvent_t e : c->seg_stage_ev)
    if (e) (void)hipEventDestroy(e);
  for (auto e : c->ev_pool) (void)hipEventDestroy(e);
  if (c->ev) (void)hipEventDestroy(c->ev);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
}

int alloc_min_state(Ctx& c) {
  if (c.m) return 0;
  const size_t n3 = 3 * (size_t)c.N, n = c.N;
  for (double** v : {&c.m, &c.g_raw, &c.g_tan, &c.g_ext, &c.g_prev, &c.y, &c.d, &c.d_prev, &c.H, &c.m_t, &c.g_t,
                     &c.r3, &c.z3, &c.p3, &c.q3, &c.cv3})
    MM_TRY(dalloc(c, v, n3));
  MM_TRY(dalloc(c, &c.U, n));
  MM_TRY(dalloc(c, &c.U_t, n));
  if (c.precision >= 1)  // the mixed forms' FP32 Steihaug vectors (else allocated by mm_set_precision)
    for (float** v : {&c.r3f, &c.z3f, &c.p3f, &c.q3f, &c.cv3f}) MM_TRY(dalloc(c, v, n3));
  MM_TRY(dalloc(c, &c.mpart, 9 * 2 * (size_t)kRedGridMax));
  MM_HIP(hipHostMalloc((void**)&c.hmpart, 9 * 2 * kRedGridMax * sizeof(double), hipHostMallocDefault));
  return 0;
}

}  // namespace mm

using namespace mm;

#define GET_CTX(h)                                  \
  Ctx* cp = get_ctx(h);                             \
  if (!cp) {                                        \
    set_error("invalid context handle");            \
    return -1;                                      \
  }                                                 \
  Ctx& c = *cp;                                     \
  MM_HIP(hipSetDevice(c.device));

extern "C" {

int mm_last_error(char* buf, int len) {
  if (buf && len > 0) {
    std::strncpy(buf, g_err.c_str(), len - 1);
    buf[len - 1] = 0;
  }
  return (int)g_err.size();
}

int mm_device_info(char* buf, int len) {
  hipDeviceProp_t prop;
  int dev = 0;
  MM_HIP(hipGetDevice(&dev));
  MM_HIP(hipGetDeviceProperties(&prop, dev));
  std::string s = std::string(prop.gcnArchName) + " CUs=" + std::to_string(prop.multiProcessorCount) +
                  " mem=" + std::to_string(prop.totalGlobalMem >> 20) + "MiB";
  if (buf && len > 0) {
    std::strncpy(buf, s.c_str(), len - 1);
    buf[len - 1] = 0;
  }
  return 0;
}

int mm_ctx_create(int N, int E, const double* knt, const int32_t* conn, const int32_t* mat_id, int G,
                  const double* A_red, const double* K1_red, const double* Js_red, const double* k_easy,
                  const double* bnd_mask, double reg, int flags, int64_t* out_ctx) {
  if (!out_ctx) {
    set_error("out_ctx is null");
    return -1;
  }
  Ctx* c = new Ctx();
  const int r = create(N, E, knt, conn, mat_id, G, A_red, K1_red, Js_red, k_easy, bnd_mask, reg, flags, *c);
  if (r != 0) {
    destroy(c);
    return r;
  }
  std::lock_guard<std::mutex> lk(g_mu);
  g_ctx.push_back(c);
  *out_ctx = (int64_t)g_ctx.size();
  return 0;
}

int mm_ctx_create_segmented(int N, int E, const double* knt, const int32_t* conn, const int32_t* mat_id, int G,
                            const double* A_red, const double* K1_red, const double* Js_red, const double* k_easy,
                            const double* bnd_mask, double reg, int flags, int nseg, const int32_t* seg_node_off,
                            int64_t* out_ctx) {
  if (!out_ctx || nseg < 1 || !seg_node_off) {
    set_error("mm_ctx_create_segmented: out_ctx, nseg >= 1 and seg_node_off are required");
    return -1;
  }
  Ctx* c = new Ctx();
  const int r = create(N, E, knt, conn, mat_id, G, A_red, K1_red, Js_red, k_easy, bnd_mask, reg, flags, *c, nseg,
                       seg_node_off);
  if (r != 0) {
    destroy(c);
    return r;
  }
  std::lock_guard<std::mutex> lk(g_mu);
  g_ctx.push_back(c);
  *out_ctx = (int64_t)g_ctx.size();
  return 0;
}

int mm_ctx_destroy(int64_t h) {
  Ctx* c = get_ctx(h);
  if (!c) {
    set_error("invalid context handle");
    return -1;
  }
  {
    std::lock_guard<std::mutex> lk(g_mu);
    g_ctx[h - 1] = nullptr;
  }
  destroy(c);
  return 0;
}

int mm_ctx_info(int64_t h, double* info) {
  GET_CTX(h);
  info[0] = c.V_mag;
  info[1] = c.N;
  info[2] = c.E;
  info[3] = c.part_all.nblk;
  info[4] = c.part_mag.nblk;
  info[5] = c.max_block_nodes;
  info[6] = c.n_mag;
  info[7] = c.nodes_all;
  info[8] = c.part_cl.nblk;
  info[9] = c.nodes_cl;
  info[10] = c.part_cl.E;
  info[11] = c.node_cap;
  info[12] = c.cl_cap;
  info[13] = c.cl_records ? 1 : 0;
  info[14] = c.apply_seq ? (double)c.h_colour_off.size() - 1 : 0.0;
  return 0;
}

int mm_ctx_get_geometry(int64_t h, int32_t* conn, double* Ve, double* grad) {
  GET_CTX(h);
  if (conn) std::memcpy(conn, c.h_conn.data(), c.h_conn.size() * sizeof(int));
  if (Ve) std::memcpy(Ve, c.h_Ve.data(), c.h_Ve.size() * sizeof(double));
  if (grad) std::memcpy(grad, c.h_grad.data(), c.h_grad.size() * sizeof(double));
  return 0;
}

int mm_ctx_get_nodal(int64_t h, double* M_nodal, double* inv_M_rel, double* inv_M_prec) {
  GET_CTX(h);
  if (M_nodal) std::memcpy(M_nodal, c.h_M_nodal.data(), c.N * sizeof(double));
  if (inv_M_rel) std::memcpy(inv_M_rel, c.h_inv_M_rel.data(), c.N * sizeof(double));
  if (inv_M_prec) std::memcpy(inv_M_prec, c.h_inv_M_prec.data(), c.N * sizeof(double));
  return 0;
}

int mm_ctx_get_exchange_diag(int64_t h, double* d) {
  GET_CTX(h);
  MM_CHECK(d, "null array");
  std::memcpy(d, c.h_exdiag.data(), c.N * sizeof(double));
  return 0;
}

int mm_set_precision(int64_t h, int precision) {
  GET_CTX(h);
  MM_CHECK(precision >= 0 && precision <= 3,
           "precision must be 0 (FP64), 1 (FP32 preconditioner), 2 (FP32 PCG with FP64 refinement) or 3 "
           "(2 with FP32 energy densities)");
  if (precision >= 2 && c.rxf == nullptr)
    for (float** v : {&c.rxf, &c.rrf, &c.rpf, &c.rqf, &c.rzf}) MM_TRY(dalloc(c, v, (size_t)c.N));
  if (precision >= 1 && c.r3f == nullptr && c.r3 != nullptr)  // contexts with minimiser state (else alloc_min_state)
    for (float** v : {&c.r3f, &c.z3f, &c.p3f, &c.q3f, &c.cv3f}) MM_TRY(dalloc(c, v, 3 * (size_t)c.N));
  c.precision = precision;
  return 0;
}

int mm_set_geometry_mode(int64_t h, int mode) {
  GET_CTX(h);
  MM_CHECK(mode >= 0 && mode <= 2, "geometry mode must be 0, 1 or 2");
  c.geom_mode = mode;
  return 0;
}

int mm_set_cluster_block_nodes(int64_t h, int nb) {
  GET_CTX(h);
  MM_CHECK(nb == 256 || nb == 512 || nb == 1024 || nb == 2048, "cluster block nodes must be 256, 512, 1024 or 2048");
  MM_CHECK(c.nseg == 1, "segmented contexts keep 256-node cluster blocks");
  if (nb == c.cl_cap) return 0;
  MM_HIP(hipStreamSynchronize(c.stream));
  HostCPart hc;
  MM_TRY(upload_cluster_part(c, nb, std::vector<int>(c.N, 0), hc, nullptr));
  if (c.apply_seq) MM_TRY(build_block_colours(c));
  if (nb != kMaxBlockNodes && c.apply_persist > 0) {
    c.apply_persist = 0;  // the persistent grid stages 256-node blocks
    c.elem_threads = 256;
  }
  if (nb != kMaxBlockNodes) c.elem_threads = 256;
  return 0;
}

int mm_set_cluster_staging(int64_t h, int mode) {
  GET_CTX(h);
  MM_CHECK(mode == 0 || mode == 1, "cluster staging must be 0 (gathers) or 1 (entry records)");
  c.cl_records = mode == 1;
  return 0;
}

int mm_set_refinement(int64_t h, double eta) {
  GET_CTX(h);
  MM_CHECK(eta > 0.0 && eta < 1.0, "residual replacement factor must be in (0, 1)");
  c.ref_eta = eta;
  return 0;
}

int mm_set_apply_scatter(int64_t h, int mode) {
  GET_CTX(h);
  MM_CHECK(mode == 0 || mode == 1, "scatter must be 0 (FP64 atomics) or 1 (colour-sequenced)");
  MM_CHECK(mode == 0 || c.nseg == 1, "segmented contexts keep the atomic scatter");
  c.apply_seq = mode == 1;
  if (c.apply_seq) MM_TRY(build_block_colours(c));
  return 0;
}

int mm_set_apply_shape(int64_t h, int unit, int param) {
  GET_CTX(h);
  MM_CHECK(unit >= 0 && unit <= 2, "unit must be 0 (element), 1 (cluster) or 2 (persistent cluster)");
  if (unit == 2) {
    MM_CHECK(param >= 1 && param <= 4, "persistent cluster apply: 1..4 workgroups per CU");
    MM_CHECK(c.cl_cap == kMaxBlockNodes, "the persistent cluster apply needs 256-node cluster blocks");
    c.apply_clusters = true;
    c.apply_persist = param;
    c.elem_threads = kMaxBlockNodes;
    return 0;
  }
  MM_CHECK(param == 64 || param == 128 || param == 256, "threads must be 64, 128 or 256");
  MM_CHECK(unit == 0 || c.cl_cap == kMaxBlockNodes || param == 256,
           "cluster blocks of more than 256 nodes run 256 threads per workgroup");
  c.apply_clusters = unit == 1;
  c.apply_persist = 0;
  c.elem_threads = param;
  return 0;
}

int mm_poisson_apply(int64_t h, const double* x, double* y) {
  GET_CTX(h);
  MM_CHECK(x && y, "null array");
  MM_HIP(hipMemcpyAsync(c.pp, x, c.N * sizeof(double), hipMemcpyHostToDevice, c.stream));
  vfill(c.pq, 0.0, c.N, c.stream);
  MM_TRY(launch_apply(c, c.geom_mode, c.pp, c.pq, c.stream));
  MM_HIP(hipGetLastError());
  MM_HIP(hipMemcpyAsync(y, c.pq, c.N * sizeof(double), hipMemcpyDeviceToHost, c.stream));
  MM_HIP(hipStreamSynchronize(c.stream));
  c.n_apply++;
  return 0;
}

int mm_poisson_apply_multi(int64_t h, int k, const double* x, double* y) {
  GET_CTX(h);
  MM_CHECK(x && y, "null array");
  MM_CHECK(k == 1 || k == 2 || k == 4 || k == 8, "k must be 1, 2, 4 or 8");
  MM_CHECK(c.geom_mode == kGeomCoords && c.apply_clusters && c.cl_cap == kMaxBlockNodes,
           "the multi-vector apply needs the on-the-fly cluster apply on 256-node blocks");
  if (!c.mx) {
    MM_TRY(dalloc(c, &c.mx, 8 * (size_t)c.N));
    MM_TRY(dalloc(c, &c.my, 8 * (size_t)c.N));
  }
  const long N = c.N;
  std::vector<double> t((size_t)k * N);
  for (int r = 0; r < k; ++r)
    for (long i = 0; i < N; ++i) t[(size_t)k * i + r] = x[(size_t)r * N + i];
  MM_HIP(hipMemcpyAsync(c.mx, t.data(), t.size() * sizeof(double), hipMemcpyHostToDevice, c.stream));
  vfill(c.my, 0.0, (long)k * N, c.stream);
  MM_TRY(launch_apply_multi(c, k, c.mx, c.my, nullptr, c.stream));
  MM_HIP(hipGetLastError());
  MM_HIP(hipMemcpyAsync(t.data(), c.my, t.size() * sizeof(double), hipMemcpyDeviceToHost, c.stream));
  MM_HIP(hipStreamSynchronize(c.stream));
  for (int r = 0; r < k; ++r)
    for (long i = 0; i < N; ++i) y[(size_t)r * N + i] = t[(size_t)k * i + r];
  c.n_apply += k;
  return 0;
}

int mm_poisson_rhs(int64_t h, const double* m, double* b) {
  GET_CTX(h);
  MM_CHECK(m && b, "null array");
  MM_TRY(alloc_min_state(c));
  MM_HIP(hipMemcpyAsync(c.m_t, m, 3 * c.N * sizeof(double), hipMemcpyHostToDevice, c.stream));
  vfill(c.pb, 0.0, c.N, c.stream);
  MM_TRY(launch_rhs(c, c.m_t, c.pb, c.stream));
  MM_HIP(hipGetLastError());
  MM_HIP(hipMemcpyAsync(b, c.pb, c.N * sizeof(double), hipMemcpyDeviceToHost, c.stream));
  MM_HIP(hipStreamSynchronize(c.stream));
  return 0;
}

int mm_poisson_diag(int64_t h, double* d) {
  GET_CTX(h);
  vfill(c.diagA, 0.0, c.N, c.stream);
  MM_TRY(launch_diag(c, c.diagA, c.stream));
  MM_HIP(hipGetLastError());
  if (d) MM_HIP(hipMemcpyAsync(d, c.diagA, c.N * sizeof(double), hipMemcpyDeviceToHost, c.stream));
  MM_HIP(hipStreamSynchronize(c.stream));
  return 0;
}

// CSR structure of a rows x cols matrix: offsets from 0, non-decreasing, column indices in range
// (the SpMV kernels index the vectors with them unchecked)
static bool valid_csr(const int32_t* p, const int32_t* j, int rows, int cols) {
  if (!p || !j || p[0] != 0) return false;
  for (int r = 0; r < rows; ++r) {
    if (p[r + 1] < p[r]) return false;
    for (int k = p[r]; k < p[r + 1]; ++k)
      if (j[k] < 0 || j[k] >= cols) return false;
  }
  return true;
}

int mm_amg_set_level(int64_t h, int l, int n, int n_next, const int32_t* Ap, const int32_t* Ai, const double* Ax,
                     const double* M, const int32_t* Pp, const int32_t* Pi, const double* Px, const int32_t* Rp,
                     const int32_t* Ri, const double* Rx, const double* dense_inv) {
  GET_CTX(h);
  MM_CHECK(l >= 0 && l < 32, "level index out of range");
  MM_CHECK(n > 0 && M, "level needs n > 0 and the smoother diagonal");
  MM_CHECK(l != 0 || n == c.N, "level 0 must have N rows");
  MM_CHECK(n_next >= 0, "n_next must be >= 0");
  if (l > 0 && Ap) MM_CHECK(Ax && valid_csr(Ap, Ai, n, n), "level matrix A: malformed CSR (n x n)");
  if (n_next > 0) {
    MM_CHECK(Pp && Pi && Px && Rp && Ri && Rx, "P and R are required above the last level");
    MM_CHECK(valid_csr(Pp, Pi, n, n_next), "prolongation P: malformed CSR (n x n_next)");
    MM_CHECK(valid_csr(Rp, Ri, n_next, n), "restriction R: malformed CSR (n_next x n)");
  }
  if ((int)c.lv.size() <= l) c.lv.resize(l + 1);
  Level L;
  L.n = n;
  L.n_next = n_next;
  MM_TRY(upload(c, &L.M, M, n));
  if (l > 0 && Ap) {
    MM_TRY(upload(c, &L.Ap, Ap, n + 1));
    MM_TRY(upload(c, &L.Ai, Ai, Ap[n]));
    MM_TRY(upload(c, &L.Ax, Ax, Ap[n]));
  }
  if (n_next > 0) {
    MM_TRY(upload(c, &L.Pp, Pp, n + 1));
    MM_TRY(upload(c, &L.Pi, Pi, Pp[n]));
    MM_TRY(upload(c, &L.Px, Px, Pp[n]));
    MM_TRY(upload(c, &L.Rp, Rp, n_next + 1));
    MM_TRY(upload(c, &L.Ri, Ri, Rp[n_next]));
    MM_TRY(upload(c, &L.Rx, Rx, Rp[n_next]));
  }
  std::vector<double> tiles64;
  if (dense_inv) {
    // lower-triangle tiles of the (symmetric) inverse; padding rows / columns are zero
    const int T = (n + 63) / 64;
    L.tiles = T;
    const size_t nt = (size_t)T * (T + 1) / 2;
    tiles64.assign(nt * 4096, 0.0);
    for (int I = 0; I < T; ++I)
      for (int J = 0; J <= I; ++J) {
        double* tile = &tiles64[((size_t)I * (I + 1) / 2 + J) * 4096];
        for (int r = 0; r < 64; ++r) {
          const int row = I * 64 + r;
          if (row >= n) break;
          for (int q = 0; q < 64; ++q) {
            const int col = J * 64 + q;
            if (col >= n) break;
            tile[r * 64 + q] = dense_inv[(size_t)row * n + col];
          }
        }
      }
    MM_TRY(upload(c, &L.dense_inv, tiles64.data(), tiles64.size()));
    MM_TRY(dalloc(c, &L.symv_part, (size_t)T * T * 64));
  }
  // FP32 copies for the mixed-precision V-cycle
  auto to_f32 = [](const double* a, size_t k) { return std::vector<float>(a, a + k); };
  MM_TRY(upload(c, &L.Mf, to_f32(M, n).data(), n));
  if (l > 0 && Ap) MM_TRY(upload(c, &L.Axf, to_f32(Ax, Ap[n]).data(), Ap[n]));
  if (n_next > 0) {
    MM_TRY(upload(c, &L.Pxf, to_f32(Px, Pp[n]).data(), Pp[n]));
    MM_TRY(upload(c, &L.Rxf, to_f32(Rx, Rp[n_next]).data(), Rp[n_next]));
  }
  if (dense_inv) MM_TRY(upload(c, &L.dense_invf_t, to_f32(tiles64.data(), tiles64.size()).data(), tiles64.size()));
  // FP16 copies for the FP32 PCG's V-cycle, when every value is finite: each matrix is scaled by
  // 2^e2 so that its largest magnitude lands in [2^13, 2^14) (below the FP16 maximum 65504, and
  // as far above the FP16 subnormals as it goes), whatever the mesh's length unit; *inv = 2^-e2.
  auto up_f16 = [&](__half** dst, float* inv, const double* a, size_t k) -> int {
    double amax = 0.0;
    for (size_t i = 0; i < k; ++i) {
      if (!std::isfinite(a[i])) return 0;
      amax = std::max(amax, std::fabs(a[i]));
    }
    int e = 0;
    if (amax > 0.0) std::frexp(amax, &e);  // amax in [2^(e-1), 2^e)
    const int e2 = 14 - e;
    if (e2 < -100 || e2 > 100) return 0;
    const double sc = std::ldexp(1.0, e2);
    std::vector<__half> h(k);
    for (size_t i = 0; i < k; ++i) h[i] = __float2half((float)(a[i] * sc));
    *inv = std::ldexp(1.0f, -e2);
    return upload(c, dst, h.data(), k);
  };
  if (l > 0 && Ap) MM_TRY(up_f16(&L.Axh, &L.Axh_s, Ax, Ap[n]));
  if (n_next > 0) {
    MM_TRY(up_f16(&L.Pxh, &L.Pxh_s, Px, Pp[n]));
    MM_TRY(up_f16(&L.Rxh, &L.Rxh_s, Rx, Rp[n_next]));
  }
  // SpMV lanes per row: power of two near a quarter of the average row length, in [4, 64]
  auto lanes = [](long nnz, long rows) {
    const double avg = rows > 0 ? (double)nnz / (double)rows : 0.0;
    int l = 4;
    while (l < 64 && 4.0 * l <= avg) l *= 2;
    return l;
  };
  if (l > 0 && Ap) L.lanesA = lanes(Ap[n], n);
  if (n_next > 0) {
    L.lanesP = lanes(Pp[n], n);
    L.lanesR = lanes(Rp[n_next], n_next);
  }
  // rows of each segment on this level (contiguous when the coarse unknowns follow the node
  // order, see amg.py), and from P those of the next level
  if ((int)c.amg_seg_off.size() <= l + 1) c.amg_seg_off.resize(l + 2);
  if (l == 0) c.amg_seg_off[0] = c.nseg > 1 ? c.h_seg_node_off : std::vector<int>();
  L.seg_off = c.amg_seg_off[l];
  c.amg_seg_off[l + 1].clear();
  if (n_next > 0 && (int)L.seg_off.size() == c.nseg + 1) {
    std::vector<int> nx(c.nseg + 1, 0);
    bool ok = true;
    int expect = 0;
    for (int g = 0; g < c.nseg && ok; ++g) {
      int lo = INT32_MAX, hi = -1;
      for (int r = L.seg_off[g]; r < L.seg_off[g + 1]; ++r)
        for (int k = Pp[r]; k < Pp[r + 1]; ++k) {
          lo = std::min(lo, Pi[k]);
          hi = std::max(hi, Pi[k]);
        }
      ok = hi >= lo && lo == expect;  // non-empty and right after the previous segment
      nx[g] = lo;
      expect = hi + 1;
    }
    ok = ok && expect == n_next;
    nx[c.nseg] = n_next;
    if (ok) c.amg_seg_off[l + 1] = nx;
  }
  if (l > 0) {
    MM_TRY(dalloc(c, &L.b, n));
    MM_TRY(dalloc(c, &L.x, n));
    MM_TRY(dalloc(c, &L.t, n));
    MM_TRY(dalloc(c, &L.xs, n));
    MM_TRY(dalloc(c, &L.bf, n));
    MM_TRY(dalloc(c, &L.xf, n));
    MM_TRY(dalloc(c, &L.tf, n));
    MM_TRY(dalloc(c, &L.xsf, n));
  }
  c.lv[l] = L;
  return 0;
}

int mm_amg_set_lanes(int64_t h, int l, int which, int lanes) {
  GET_CTX(h);
  MM_CHECK(l >= 0 && l < (int)c.lv.size(), "level index out of range");
  MM_CHECK(which >= 0 && which <= 2, "which must be 0 (A), 1 (P) or 2 (R)");
  MM_CHECK(lanes == 1 || lanes == 2 || lanes == 4 || lanes == 8 || lanes == 16 || lanes == 32 || lanes == 64,
           "lanes per row must be 1, 2, 4, 8, 16, 32 or 64");
  Level& L = c.lv[l];
  (which == 0 ? L.lanesA : which == 1 ? L.lanesP : L.lanesR) = lanes;
  return 0;
}

int mm_amg_get_lanes(int64_t h, int l, int32_t* out3) {
  GET_CTX(h);
  MM_CHECK(l >= 0 && l < (int)c.lv.size() && out3, "level index out of range");
  out3[0] = c.lv[l].lanesA;
  out3[1] = c.lv[l].lanesP;
  out3[2] = c.lv[l].lanesR;
  return 0;
}

int mm_amg_commit(int64_t h, int kind, int nlev) {
  GET_CTX(h);
  c.blist_key.clear();  // segment subsets recompute the row runs of the new levels
  MM_CHECK(nlev >= 1 && (int)c.lv.size() >= nlev, "levels missing");
  c.lv.resize(nlev);
  for (int l = 0; l < nlev; ++l) {
    MM_CHECK(c.lv[l].M != nullptr, "level not set");
    if (l < nlev - 1) {
      MM_CHECK(c.lv[l].n_next == c.lv[l + 1].n, "P/R shape does not match the next level");
      MM_CHECK(l == 0 || c.lv[l].Ap, "coarse level needs A");
    } else {
      MM_CHECK(c.lv[l].dense_inv != nullptr, "last level needs the dense inverse");
    }
  }
  c.amg_kind = kind;
  c.lv_half = nlev >= 2;
  for (int l = 0; l + 1 < nlev; ++l)
    c.lv_half = c.lv_half && c.lv[l].Pxh && c.lv[l].Rxh && (l == 0 || c.lv[l].Axh);
  return 0;
}

int mm_precond_jacobi(int64_t h) {
  GET_CTX(h);
  MM_TRY(mm_poisson_diag(h, nullptr));
  std::vector<double> d(c.N);
  MM_HIP(hipMemcpyAsync(d.data(), c.diagA, c.N * sizeof(double), hipMemcpyDeviceToHost, c.stream));
  MM_HIP(hipStreamSynchronize(c.stream));
  for (auto& v : d) v = 1.0 / (v + 1e-30);
  c.lv.clear();
  c.lv.resize(1);
  c.lv[0].n = c.N;
  MM_TRY(upload(c, &c.lv[0].M, d.data(), c.N));
  c.amg_kind = 0;
  c.lv_half = false;
  return 0;
}

int mm_precond_apply(int64_t h, const double* r, double* z) {
  GET_CTX(h);
  MM_CHECK(r && z, "null array");
  MM_HIP(hipMemcpyAsync(c.pr, r, c.N * sizeof(double), hipMemcpyHostToDevice, c.stream));
  MM_TRY(precond_apply(c, c.pr, c.pz, c.part_rz, c.stream));
  MM_HIP(hipGetLastError());
  MM_HIP(hipMemcpyAsync(z, c.pz, c.N * sizeof(double), hipMemcpyDeviceToHost, c.stream));
  MM_HIP(hipStreamSynchronize(c.stream));
  return 0;
}

int mm_pcg_solve(int64_t h, const double* b, double* x_inout, double tol, int maxiter, int* it, double* r2) {
  GET_CTX(h);
  MM_CHECK(b && x_inout, "null array");
  MM_HIP(hipMemcpyAsync(c.pb, b, c.N * sizeof(double), hipMemcpyHostToDevice, c.stream));
  MM_HIP(hipMemcpyAsync(c.work_U, x_inout, c.N * sizeof(double), hipMemcpyHostToDevice, c.stream));
  int iters = 0;
  double rr = 0, bb = 0;
  MM_TRY(pcg_dev(c, c.pb, c.work_U, tol, maxiter, &iters, &rr, &bb));
  MM_HIP(hipGetLastError());
  MM_HIP(hipMemcpyAsync(x_inout, c.work_U, c.N * sizeof(double), hipMemcpyDeviceToHost, c.stream));
  MM_HIP(hipStreamSynchronize(c.stream));
  if (it) *it = iters;
  if (r2) *r2 = rr;
  return 0;
}

int mm_solve_U(int64_t h, const double* m, double* U_inout, double tol, int maxiter, int* it, double* rel_res) {
  GET_CTX(h);
  MM_CHECK(m && U_inout, "null array");
  MM_TRY(alloc_min_state(c));
  MM_HIP(hipMemcpyAsync(c.m_t, m, 3 * c.N * sizeof(double), hipMemcpyHostToDevice, c.stream));
  MM_HIP(hipMemcpyAsync(c.work_U, U_inout, c.N * sizeof(double), hipMemcpyHostToDevice, c.stream));
  int iters = 0;
  double rr = 0;
  MM_TRY(solve_U_dev(c, c.m_t, c.work_U, tol, maxiter, &iters, &rr));
  MM_HIP(hipGetLastError());
  MM_HIP(hipMemcpyAsync(U_inout, c.work_U, c.N * sizeof(double), hipMemcpyDeviceToHost, c.stream));
  MM_HIP(hipStreamSynchronize(c.stream));
  if (it) *it = iters;
  if (rel_res) *rel_res = rr;
  return 0;
}

int mm_solve_U_segmented(int64_t h, const double* m, double* U_inout, const double* tol, int maxiter, int* it,
                         double* rel_res) {
  GET_CTX(h);
  MM_CHECK(m && U_inout && tol && it && rel_res, "null argument");
  MM_CHECK(maxiter >= 0, "maxiter must be >= 0");
  MM_TRY(alloc_min_state(c));
  MM_HIP(hipMemcpyAsync(c.m_t, m, 3 * c.N * sizeof(double), hipMemcpyHostToDevice, c.stream));
  MM_HIP(hipMemcpyAsync(c.work_U, U_inout, c.N * sizeof(double), hipMemcpyHostToDevice, c.stream));
  MM_TRY(solve_U_seg(c, c.m_t, c.work_U, tol, maxiter, it, rel_res));
  MM_HIP(hipGetLastError());
  MM_HIP(hipMemcpyAsync(U_inout, c.work_U, c.N * sizeof(double), hipMemcpyDeviceToHost, c.stream));
  MM_HIP(hipStreamSynchronize(c.stream));
  return 0;
}

int mm_energy_grad(int64_t h, const double* m, const double* U, const double* B, double* E, double* g) {
  GET_CTX(h);
  MM_CHECK(m && U && B && E && g, "null array");
  MM_TRY(alloc_min_state(c));
  MM_HIP(hipMemcpyAsync(c.m_t, m, 3 * c.N * sizeof(double), hipMemcpyHostToDevice, c.stream));
  MM_HIP(hipMemcpyAsync(c.U_t, U, c.N * sizeof(double), hipMemcpyHostToDevice, c.stream));
  vfill(c.g_t, 0.0, 3L * c.N, c.stream);
  MM_TRY(launch_energy(c, c.m_t, c.U_t, B, c.g_t, c.dscal + 16, c.stream));
  MM_HIP(hipGetLastError());
  MM_HIP(hipMemcpyAsync(g, c.g_t, 3 * c.N * sizeof(double), hipMemcpyDeviceToHost, c.stream));
  MM_HIP(hipMemcpyAsync(E, c.dscal + 16, sizeof(double), hipMemcpyDeviceToHost, c.stream));
  MM_HIP(hipStreamSynchronize(c.stream));
  c.n_energy++;
  return 0;
}

int mm_energy_grad_noalloc(int64_t h, const double* m_dev, const double* U_dev, const double* B, double* E_dev,
                           double* g_dev, void* stream) {
  GET_CTX(h);
  MM_CHECK(m_dev && U_dev && B && E_dev && g_dev, "null argument");
  MM_CHECK(2L * c.part_mag.nblk <= kScratchDoubles,
           "mesh too large for the allocation-free energy: 2 x magnetic blocks exceeds the 65536-double scratch");
  double* sp = scratch_on(c.device);
  MM_CHECK(sp != nullptr, "scratch address not resolved for the context's device");
  hipStream_t s = stream ? (hipStream_t)stream : c.stream;
  vfill(g_dev, 0.0, 3L * c.N, s);
  MM_TRY(launch_energy_buf(c, m_dev, U_dev, B, g_dev, E_dev, sp, s));
  MM_HIP(hipGetLastError());
  c.n_energy++;
  return 0;
}

int mm_local_grad(int64_t h, const double* v, double* g) {
  GET_CTX(h);
  MM_CHECK(v && g, "null array");
  MM_TRY(alloc_min_state(c));
  MM_HIP(hipMemcpyAsync(c.p3, v, 3 * c.N * sizeof(double), hipMemcpyHostToDevice, c.stream));
  vfill(c.cv3, 0.0, 3L * c.N, c.stream);
  MM_TRY(launch_local(c, c.p3, c.cv3, nullptr, c.stream));
  MM_HIP(hipGetLastError());
  MM_HIP(hipMemcpyAsync(g, c.cv3, 3 * c.N * sizeof(double), hipMemcpyDeviceToHost, c.stream));
  MM_HIP(hipStreamSynchronize(c.stream));
  c.n_local++;
  return 0;
}

int mm_j_parallel(int64_t h, const double* m, const double* hdir, double* J) {
  GET_CTX(h);
  MM_CHECK(m && hdir && J, "null argument");
  MM_TRY(alloc_min_state(c));
  MM_HIP(hipMemcpyAsync(c.m_t, m, 3 * c.N * sizeof(double), hipMemcpyHostToDevice, c.stream));
  double hn = std::sqrt(hdir[0] * hdir[0] + hdir[1] * hdir[1] + hdir[2] * hdir[2]) + 1e-30;
  const double hh[3] = {hdir[0] / hn, hdir[1] / hn, hdir[2] / hn};
  MM_TRY(launch_jpar(c, c.m_t, hh, c.dscal + 20, c.stream));
  MM_HIP(hipMemcpyAsync(c.hscal + 20, c.dscal + 20, 2 * sizeof(double), hipMemcpyDeviceToHost, c.stream));
  MM_HIP(hipStreamSynchronize(c.stream));
  *J = c.hscal[20] / (c.hscal[21] + 1e-30);
  return 0;
}

int mm_state_j_parallel(int64_t h, const double* hdir, double* J) {
  GET_CTX(h);
  MM_CHECK(hdir && J, "null argument");
  MM_CHECK(c.have_state, "no resident state");
  double hn = std::sqrt(hdir[0] * hdir[0] + hdir[1] * hdir[1] + hdir[2] * hdir[2]) + 1e-30;
  const double hh[3] = {hdir[0] / hn, hdir[1] / hn, hdir[2] / hn};
  MM_TRY(launch_jpar(c, c.m, hh, c.dscal + 20, c.stream));
  MM_HIP(hipMemcpyAsync(c.hscal + 20, c.dscal + 20, 2 * sizeof(double), hipMemcpyDeviceToHost, c.stream));
  MM_HIP(hipStreamSynchronize(c.stream));
  *J = c.hscal[20] / (c.hscal[21] + 1e-30);
  return 0;
}

int mm_state_set(int64_t h, const double* m, const double* U) {
  GET_CTX(h);
  MM_CHECK(m, "null argument");
  MM_TRY(alloc_min_state(c));
  MM_HIP(hipMemcpyAsync(c.m, m, 3 * c.N * sizeof(double), hipMemcpyHostToDevice, c.stream));
  if (U)
    MM_HIP(hipMemcpyAsync(c.U, U, c.N * sizeof(double), hipMemcpyHostToDevice, c.stream));
  else
    vfill(c.U, 0.0, c.N, c.stream);
  MM_HIP(hipStreamSynchronize(c.stream));
  c.have_state = true;
  return 0;
}

int mm_state_get(int64_t h, double* m, double* U) {
  GET_CTX(h);
  MM_CHECK(c.have_state, "no resident state");
  if (m) MM_HIP(hipMemcpyAsync(m, c.m, 3 * c.N * sizeof(double), hipMemcpyDeviceToHost, c.stream));
  if (U) MM_HIP(hipMemcpyAsync(U, c.U, c.N * sizeof(double), hipMemcpyDeviceToHost, c.stream));
  MM_HIP(hipStreamSynchronize(c.stream));
  return 0;
}

static void to_minparams(const mm_params* p, MinParams& q) {
  q.method = p->method;
  q.max_iter = p->max_iter;
  q.tau_f = p->tau_f;
  q.eps_a = p->eps_a;
  q.tau0 = p->tau0;
  q.ls_eta1 = p->ls_eta1;
  q.ls_eta2 = p->ls_eta2;
  q.ls_C = p->ls_C;
  q.ls_c = p->ls_c;
  q.ls_max_evals = p->ls_max_evals;
  q.pc_iters = p->pc_iters;
  q.pc_auto = p->pc_auto;
  q.pc_force_eta = p->pc_force_eta;
  q.pc_force_alpha = p->pc_force_alpha;
  q.pc_stagnation_nu = p->pc_stagnation_nu;
  q.pc_reg = p->pc_reg;
  q.L = p->L;
  q.cg_tol = p->cg_tol;
  q.phi_tol = p->phi_tol;
  q.cg_maxiter = p->cg_maxiter;
  q.memory = p->memory;
}

// Parameter checks shared by the minimiser entry points. The L-BFGS history keeps its alphas and
// rhos in device scalar slots S_ALPHAS.. and S_RHOS.. (32 each), so memory is capped at 32.
static int check_params(const mm_params* p) {
  MM_CHECK(p, "null parameter struct");
  MM_CHECK(p->method == 0 || p->method == 1, "method must be 0 (pcohen_hs) or 1 (lbfgs)");
  MM_CHECK(p->memory >= 0 && p->memory <= kMaxLbfgsMemory, "memory (L-BFGS history) must be in [1, 32] (0: 10)");
  MM_CHECK(p->max_iter >= 0 && p->ls_max_evals >= 1 && p->pc_iters >= 0 && p->cg_maxiter >= 1 && p->L >= 0,
           "iteration limits must be non-negative (ls_max_evals, cg_maxiter >= 1)");
  for (double v : {p->tau_f, p->eps_a, p->tau0, p->ls_eta1, p->ls_eta2, p->ls_C, p->ls_c, p->pc_force_eta,
                   p->pc_force_alpha, p->pc_stagnation_nu, p->pc_reg, p->cg_tol, p->phi_tol})
    MM_CHECK(std::isfinite(v) && v >= 0.0, "tolerances and line-search constants must be finite and non-negative");
  return 0;
}

int mm_params_check(const mm_params* p) { return check_params(p); }

int mm_minimize_resident(int64_t h, const mm_params* p, const double* B, mm_info* info) {
  GET_CTX(h);
  MM_CHECK(p && B && info, "null argument");
  MM_CHECK(c.have_state, "no resident state (call mm_state_set)");
  MM_TRY(check_params(p));
  MM_CHECK(c.nseg == 1, "the context holds several independent systems: use mm_minimize_segmented");
  MinParams q;
  to_minparams(p, q);
  MinInfo mi;
  MM_TRY(minimize_resident(c, q, B, mi));
  info->iters = mi.iters;
  info->evals = mi.evals;
  info->demag_iters = mi.demag_iters;
  info->preco_iters = mi.preco_iters;
  info->converged = mi.converged;
  info->E = mi.E;
  info->gnorm = mi.gnorm;
  info->time = mi.time;
  return 0;
}

int mm_set_solve_guess(int64_t h, int k) {
  GET_CTX(h);
  return set_solve_guess(c, k);
}

int mm_minimize_segmented(int64_t h, const mm_params* p, const double* B, mm_info* info) {
  GET_CTX(h);
  MM_CHECK(p && B && info, "null argument");
  MM_CHECK(c.have_state, "no resident state (call mm_state_set)");
  MM_CHECK(c.nseg > 1, "mm_minimize_segmented needs a context created with mm_ctx_create_segmented");
  MM_TRY(check_params(p));
  MM_CHECK(p->method == 0, "segmented minimisation runs pcohen_hs (method 0)");
  MM_CHECK(c.lv.size() > 1 && c.lv[0].M != nullptr && c.geom_mode == kGeomCoords,
           "segmented minimisation needs the AMG V-cycle and the on-the-fly apply");
  MinParams q;
  to_minparams(p, q);
  std::vector<MinInfo> mi(c.nseg);
  MM_TRY(minimize_segmented(c, q, B, mi.data()));
  for (int g = 0; g < c.nseg; ++g) {
    info[g].iters = mi[g].iters;
    info[g].evals = mi[g].evals;
    info[g].demag_iters = mi[g].demag_iters;
    info[g].preco_iters = mi[g].preco_iters;
    info[g].converged = mi[g].converged;
    info[g].E = mi[g].E;
    info[g].gnorm = mi[g].gnorm;
    info[g].time = mi[g].time;
  }
  return 0;
}

int mm_state_j_parallel_segmented(int64_t h, const double* hdir, double* J) {
  GET_CTX(h);
  MM_CHECK(hdir && J, "null argument");
  MM_CHECK(c.have_state, "no resident state");
  MM_CHECK(c.nseg > 1, "mm_state_j_parallel_segmented needs a segmented context");
  const double hn = std::sqrt(hdir[0] * hdir[0] + hdir[1] * hdir[1] + hdir[2] * hdir[2]) + 1e-30;
  const double hh[3] = {hdir[0] / hn, hdir[1] / hn, hdir[2] / hn};
  return j_parallel_segmented(c, c.m, hh, J);
}

int mm_seg_volumes(int64_t h, double* V_mag) {
  GET_CTX(h);
  MM_CHECK(V_mag, "null argument");
  for (int g = 0; g < c.nseg; ++g) V_mag[g] = c.h_seg_vmag[g];
  return 0;
}

int mm_minimize(int64_t h, const mm_params* p, double* m_inout, double* U_inout, const double* B, mm_info* info) {
  MM_TRY(mm_state_set(h, m_inout, U_inout));
  MM_TRY(mm_minimize_resident(h, p, B, info));
  return mm_state_get(h, m_inout, U_inout);
}

}  // extern "C"

// Runs job(i) for i in [0, n) on up to n_threads host threads (each takes the next index);
// the first failing index's message becomes the caller's error.
template <typename Job>
static int run_batched(int n, const int64_t* ctxs, int n_threads, Job job) {
  MM_CHECK(n >= 0 && (n == 0 || ctxs), "null context list");
  for (int i = 0; i < n; ++i) {
    MM_CHECK(get_ctx(ctxs[i]) != nullptr, "invalid context handle in batch");
    for (int j = 0; j < i; ++j) MM_CHECK(ctxs[j] != ctxs[i], "a context appears twice in one batch");
  }
  const int T = std::max(1, std::min(n, n_threads > 0 ? n_threads : 8));
  std::atomic<int> next{0};
  std::mutex emu;
  int first_bad = n, first_rc = 0;
  std::string first_msg;
  auto worker = [&]() {
    for (int i = next++; i < n; i = next++) {
      Ctx& c = *get_ctx(ctxs[i]);
      int rc = 0;
      if (hipSetDevice(c.device) != hipSuccess) {
        set_error("hipSetDevice failed");
        rc = -1;
      } else {
        rc = job(i);
      }
      if (rc != 0) {
        std::lock_guard<std::mutex> lk(emu);
        if (i < first_bad) {
          first_bad = i;
          first_rc = rc;
          first_msg = g_err;
        }
      }
    }
  };
  if (T == 1) {
    worker();
  } else {
    std::vector<std::thread> pool;
    pool.reserve(T);
    for (int t = 0; t < T; ++t) pool.emplace_back(worker);
    for (auto& th : pool) th.join();
  }
  if (first_bad < n) {
    set_error("batch item " + std::to_string(first_bad) + ": " + first_msg);
    return first_rc;
  }
  return 0;
}

extern "C" {

int mm_minimize_batched(int n, const int64_t* ctxs, const mm_params* p, const double* B, int n_threads,
                        mm_info* out) {
  MM_CHECK(p && B && out, "null argument");
  return run_batched(n, ctxs, n_threads, [&](int i) { return mm_minimize_resident(ctxs[i], p, B + 3 * i, out + i); });
}

int mm_solve_U_batched(int n, const int64_t* ctxs, const double* const* m, double* const* U_inout, const double* tol,
                       int maxiter, int n_threads, int* it, double* rel_res) {
  MM_CHECK(m && U_inout && tol && it && rel_res, "null argument");
  return run_batched(n, ctxs, n_threads, [&](int i) {
    return mm_solve_U(ctxs[i], m[i], U_inout[i], tol[i], maxiter, it + i, rel_res + i);
  });
}

int mm_edge_clusters(int N, int E, const double* knt, const int32_t* conn, int32_t* nodes8, uint32_t* tets,
                     int32_t* n_clusters) {
  MM_CHECK(N > 0 && E > 0 && knt && conn && nodes8 && tets && n_clusters, "bad arguments");
  std::vector<int> cv(conn, conn + 4 * (size_t)E);
  for (int v : cv) MM_CHECK(v >= 0 && v < N, "connectivity index out of range");
  std::vector<HostCluster> cl;
  build_clusters(knt, E, cv, cl);
  MM_CHECK(cl.size() <= (size_t)E, "more clusters than tets");
  for (size_t i = 0; i < cl.size(); ++i) {
    for (int q = 0; q < 8; ++q) nodes8[8 * i + q] = cl[i].node[q];
    tets[i] = cl[i].tets;
  }
  *n_clusters = (int32_t)cl.size();
  return 0;
}

int mm_sa_aggregate(int n, const int32_t* Sp, const int32_t* Sj, int32_t* agg, int32_t* n_agg) {
  // Greedy three-pass aggregation on the strength graph S (CSR, diagonal ignored):
  // 1) a node whose neighbours are all free becomes a root and takes its whole neighbourhood;
  // 2) remaining free nodes join the aggregate of an aggregated neighbour;
  // 3) still-free nodes start new aggregates with their free neighbours.
  // Nodes without neighbours stay unaggregated (agg = -1).
  if (n < 0 || !Sp || !Sj || !agg || !n_agg) {
    set_error("mm_sa_aggregate: bad arguments");
    return -1;
  }
  std::vector<int> state(n, 0);  // 0 free, >0 aggregate+1 (pass 1/3), <0 joined -(aggregate+1), isolated -n-1
  int next = 0;
  const int iso = -n - 1;
  for (int i = 0; i < n; ++i) {
    if (state[i] != 0) continue;
    bool has_nb = false, nb_taken = false;
    for (int k = Sp[i]; k < Sp[i + 1]; ++k) {
      const int j = Sj[k];
      if (j == i) continue;
      has_nb = true;
      if (state[j] != 0) {
        nb_taken = true;
        break;
      }
    }
    if (!has_nb) {
      state[i] = iso;
    } else if (!nb_taken) {
      ++next;
      state[i] = next;
      for (int k = Sp[i]; k < Sp[i + 1]; ++k) state[Sj[k]] = next;
    }
  }
  for (int i = 0; i < n; ++i) {
    if (state[i] != 0) continue;
    for (int k = Sp[i]; k < Sp[i + 1]; ++k) {
      const int sj = state[Sj[k]];
      if (sj > 0) {
        state[i] = -sj;
        break;
      }
    }
  }
  int count = next;  // aggregates are numbered 0..count-1 so far
  for (int i = 0; i < n; ++i) {
    const int si = state[i];
    if (si == iso) {
      agg[i] = -1;
    } else if (si > 0) {
      agg[i] = si - 1;
    } else if (si < 0) {
      agg[i] = -si - 1;
    } else {
      agg[i] = count;
      for (int k = Sp[i]; k < Sp[i + 1]; ++k) {
        const int j = Sj[k];
        if (j != i && state[j] == 0) {
          state[j] = -(count + 1);
          if (j < i) agg[j] = count;
        }
      }
      state[i] = -(count + 1);
      ++count;
    }
  }
  *n_agg = count;
  return 0;
}

int mm_apply_timing(int64_t h, int enable, double* total_ms, int64_t* count) {
  GET_CTX(h);
  MM_TRY(resolve_apply_timing(c));
  if (total_ms) *total_ms = c.apply_ms;
  if (count) *count = c.apply_timed;
  c.apply_ms = 0.0;
  c.apply_timed = 0;
  c.apply_seen = 0;
  c.time_apply = enable != 0;
  return 0;
}

int mm_bench_precond(int64_t h, int use_graph, int reps, double* ms) {
  GET_CTX(h);
  MM_CHECK(reps > 0 && c.lv.size() >= 1 && c.lv[0].M, "bad arguments or no preconditioner");
  const bool timing = c.time_apply;
  c.time_apply = false;
  vfill(c.pr, 1.0, c.N, c.stream);
  hipEvent_t e0, e1;
  MM_HIP(hipEventCreate(&e0));
  MM_HIP(hipEventCreate(&e1));
  MM_TRY(precond_apply(c, c.pr, c.pz, c.part_rz, c.stream));  // warm-up
  hipGraphExec_t exec = nullptr;
  hipGraph_t graph = nullptr;
  if (use_graph) {
    MM_HIP(hipStreamBeginCapture(c.stream, hipStreamCaptureModeThreadLocal));
    vfill(c.pt, 0.0, c.N, c.stream);
    if (c.ptf) vfill_f(c.ptf, 0.0f, c.N, c.stream);
    MM_TRY(precond(c, c.pr, c.pz, c.part_rz, c.stream));
    MM_HIP(hipStreamEndCapture(c.stream, &graph));
    MM_HIP(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
    MM_HIP(hipGraphLaunch(exec, c.stream));
  }
  MM_HIP(hipEventRecord(e0, c.stream));
  for (int r = 0; r < reps; ++r) {
    if (use_graph) {
      MM_HIP(hipGraphLaunch(exec, c.stream));
    } else {
      vfill(c.pt, 0.0, c.N, c.stream);
      if (c.ptf) vfill_f(c.ptf, 0.0f, c.N, c.stream);
      MM_TRY(precond(c, c.pr, c.pz, c.part_rz, c.stream));
    }
  }
  MM_HIP(hipEventRecord(e1, c.stream));
  MM_HIP(hipEventSynchronize(e1));
  float t = 0;
  MM_HIP(hipEventElapsedTime(&t, e0, e1));
  *ms = t / reps;
  if (exec) (void)hipGraphExecDestroy(exec);
  if (graph) (void)hipGraphDestroy(graph);
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  c.time_apply = timing;
  return 0;
}

int mm_counters(int64_t h, int64_t* out6) {
  GET_CTX(h);
  out6[0] = c.n_apply;
  out6[1] = c.n_rhs;
  out6[2] = c.n_energy;
  out6[3] = c.n_local;
  out6[4] = c.n_vcycle;
  out6[5] = c.n_pcg_it;
  return 0;
}

int mm_bench_apply_multi(int64_t h, int k, int reps, double* ms) {
  GET_CTX(h);
  MM_CHECK(reps > 0 && ms, "bad arguments");
  MM_CHECK(k == 1 || k == 2 || k == 4 || k == 8, "k must be 1, 2, 4 or 8");
  MM_CHECK(c.apply_clusters && c.cl_cap == kMaxBlockNodes, "the multi-vector apply needs 256-node cluster blocks");
  if (!c.mx) {
    MM_TRY(dalloc(c, &c.mx, 8 * (size_t)c.N));
    MM_TRY(dalloc(c, &c.my, 8 * (size_t)c.N));
  }
  if (c.mdot_n < c.part_cl.nblk) {  // k partials per block
    MM_TRY(dalloc(c, &c.mdot, 8 * (size_t)c.part_cl.nblk));
    c.mdot_n = c.part_cl.nblk;
  }
  hipEvent_t e0, e1;
  MM_HIP(hipEventCreate(&e0));
  MM_HIP(hipEventCreate(&e1));
  vfill(c.mx, 1.0, (long)k * c.N, c.stream);
  vfill(c.my, 0.0, (long)k * c.N, c.stream);
  MM_TRY(launch_apply_multi(c, k, c.mx, c.my, c.mdot, c.stream));  // warm-up
  MM_HIP(hipEventRecord(e0, c.stream));
  for (int r = 0; r < reps; ++r) MM_TRY(launch_apply_multi(c, k, c.mx, c.my, c.mdot, c.stream));
  MM_HIP(hipEventRecord(e1, c.stream));
  MM_HIP(hipEventSynchronize(e1));
  float t = 0;
  MM_HIP(hipEventElapsedTime(&t, e0, e1));
  *ms = t / reps;
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return 0;
}

int mm_bench_copy(int64_t h, int reps, double* ms) {
  GET_CTX(h);
  MM_CHECK(reps > 0 && ms, "bad arguments");
  hipEvent_t e0, e1;
  MM_HIP(hipEventCreate(&e0));
  MM_HIP(hipEventCreate(&e1));
  vfill(c.pp, 1.0, c.N, c.stream);
  MM_HIP(hipEventRecord(e0, c.stream));
  for (int r = 0; r < reps; ++r) MM_TRY(launch_copy_calib(c.pq, c.pp, c.N, c.stream));
  MM_HIP(hipEventRecord(e1, c.stream));
  MM_HIP(hipEventSynchronize(e1));
  float t = 0;
  MM_HIP(hipEventElapsedTime(&t, e0, e1));
  *ms = t / reps;
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return 0;
}

int mm_bench_apply(int64_t h, int geom, int reps, double* ms) {
  GET_CTX(h);
  MM_CHECK(geom >= 0 && geom <= 3 && reps > 0, "bad arguments");
  hipEvent_t e0, e1;
  MM_HIP(hipEventCreate(&e0));
  MM_HIP(hipEventCreate(&e1));
  vfill(c.pp, 1.0, c.N, c.stream);
  vfill(c.pq, 0.0, c.N, c.stream);
  MM_HIP(hipMemsetAsync(c.px0f, 0, c.N * sizeof(float), c.stream));
  // geom 3: the FP32 apply of the mixed-precision preconditioner (x, y in the FP32 work vectors)
  auto one = [&]() { return geom == 3 ? launch_apply_f32(c, c.px0f, c.ptf, c.stream) : launch_apply(c, geom, c.pp, c.pq, c.stream); };
  MM_TRY(one());  // warm-up
  MM_HIP(hipEventRecord(e0, c.stream));
  for (int r = 0; r < reps; ++r) MM_TRY(one());
  MM_HIP(hipEventRecord(e1, c.stream));
  MM_HIP(hipEventSynchronize(e1));
  float t = 0;
  MM_HIP(hipEventElapsedTime(&t, e0, e1));
  *ms = t / reps;
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return 0;
}

}  // extern "C"
