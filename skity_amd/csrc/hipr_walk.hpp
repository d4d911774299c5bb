// hipr_walk.hpp -- stage (b): the analytic-AA scanline walk of one path per
// wavefront, producing 16x16 coverage tiles.
//
// This is WalkEdges (src/render/sw/sw_raster.cc:546-677) with
// insert_new_edges / backward_insert_edge_based_on_x / check_intersection
// (sw_raster.cc:159-247) re-expressed for a 64-lane wave.  The active edge
// list lives in registers (element p in lane p % 64, slot p / 64) and every
// sequential list operation of the reference is replaced by the parallel
// operation with the same result:
//   * winding sums and interval roles: wave prefix sums in list order;
//   * the left edge of an interval / the previous right edge: max-scans;
//   * the per-edge x re-insertion after a step keeps the survivors stably
//     sorted by x (backward_insert_edge_based_on_x moves an edge behind the
//     processed edges with larger x), and check_intersection compares each
//     survivor with its predecessor among the already processed survivors;
//   * insert_new_edges places the first new edge after the last active edge
//     with x <= its x and every later one before the first following active
//     edge with x >= its x; new edges arrive sorted by (upper_y, x, dx, index)
//     (SortEdges, sw_raster.cc:679-704, with emission order breaking ties).
// Coverage of a pixel row is collected as LDS deltas (hipr_blit.hpp) and, when
// the row is complete, resolved into at most two covers per pixel (the
// directly emitted span first, then the accumulated one), which are written
// into the plane(s) of the 16x16 tile holding the pixel.
#ifndef SKITY_HIP_WALK_HPP_
#define SKITY_HIP_WALK_HPP_

#include "hipr_blit.hpp"
#include "hipr_common.hpp"

namespace hipr {

constexpr int kTilePlaneBytes = kTile * kTile;
// Tile slot of an allocation that found no plane (capacity and pool spent).
constexpr int32_t kNoPlane = 0x7FFFFFFF;

// Walk modes: count the tile planes, write them, or only follow the step
// schedule (no coverage) to record the row-band snapshots.
enum WalkMode : int { kWalkCount = 0, kWalkWrite = 1, kWalkSchedule = 2 };

// Row bands.  A path taller than `band_rows` is walked by several wavefronts
// at once, each over a range of rows.  A sequential pass (count or schedule)
// records the complete walk state at the start of every band after the
// first: the first loop iteration that begins on a pixel row that is a
// multiple of 16 (a tile row boundary) at or below the band's nominal start
// row.  Restarted from that state, the walk of a band is the same computation
// as the sequential walk over those rows, and every tile row belongs to one
// band.
// Widest walk: 64 slots x 64 lanes, an active list of 4096 edges.
constexpr int kMaxSlots = 64;
constexpr int kBandBits = 5;  // work items are job << kBandBits | band
constexpr int kMaxBands = 1 << kBandBits;
struct SnapHeader {
  int32_t y, nny, pp, n;  // y = kMaxS32: the band has no rows
  int32_t pad[12];
};
static_assert(sizeof(SnapHeader) == 64, "snapshot header is 64 bytes");
// header, then X, DX, DY, LY, M [cap] and the quad edge state [cap]
__host__ __device__ inline int64_t SnapStride(int cap) {
  return (int64_t)sizeof(SnapHeader) + (int64_t)cap * (5 * 4 + (int64_t)sizeof(EdgeRec));
}

struct WalkParams {
  const Job* jobs;
  const int32_t* walk_list;  // work items: job << 4 | band
  int32_t n_walk;
  const JobGeo* geo;
  const int64_t* edge_base;
  const EdgeRec* wedges;    // per job, sorted by (upper_y, index); read only
  EdgeRec* qstate;          // quad edge state: nE records per job and band
  const int64_t* q_base;    // per job: first qstate record of band 0
  uint8_t* snaps;           // band snapshots
  const int64_t* snap_base; // per job: byte offset of band 1's snapshot
  const int32_t* band_count;  // per job: row bands (1 = one sequential walk)
  const Canvas* canvases;
  const int64_t* plane_base;  // per job: first plane of its estimated range
  const int32_t* plane_cap;
  uint8_t* planes;
  // Shared overflow pool (estimate modes): a job's planes beyond its
  // capacity come from planes [pool_base, pool_base + pool_cap), taken with
  // an atomic counter (starts at 0).  Maps hold absolute plane numbers.
  int32_t* pool_ctr;
  int32_t pool_base, pool_cap;
  const int64_t* map_base;
  int32_t* map0;
  int32_t* map1;
  JobCov* cov;
  int32_t lds_width;        // row buffer capacity in pixels
  int32_t* queue;           // persistent form: item counter (starts at 0), else null
};

// Job pixel bounds as RastePath computes them (sw_raster.cc:741-754).
// Device-space coordinates beyond +-kCoordLimit are refused before any walk
// (the reference's float -> 26.6 conversion overflows there); the clamp keeps
// every width and height of a refused job within int32.
constexpr float kCoordLimit = 8388608.f;  // 2^23 px
// Source coordinates are refused far beyond it (a canvas matrix could bring
// them back into range, so the device-space check is the one that counts).
constexpr float kSourceCoordLimit = 1073741824.f;  // 2^30
struct PixBounds {
  int l, t, r, b;
  bool empty;
};
__host__ __device__ inline float ClampCoord(float v) {
  return v < -kCoordLimit ? -kCoordLimit : (v > kCoordLimit ? kCoordLimit : v);
}
__host__ __device__ inline PixBounds JobPixBounds(const JobGeo& g) {
  PixBounds pb;
  bool ok = g.l < g.r && g.t < g.b;
  float l = ok ? floorf(ClampCoord(g.l)) : 0.f, t = ok ? floorf(ClampCoord(g.t)) : 0.f;
  float r = ok ? ceilf(ClampCoord(g.r)) : 0.f, b = ok ? ceilf(ClampCoord(g.b)) : 0.f;
  pb.empty = !(l < r && t < b);
  pb.l = (int)l;
  pb.t = (int)t;
  pb.r = (int)r;
  pb.b = (int)b;
  return pb;
}
// Scan bounds of a job (sw_raster.cc:741-754): the quad path's bounds
// intersected with the job's clip bounds (Rect::Intersect, rect.cc:160-172:
// empty unless l < r and t < b), floored / ceiled.  Without a clip they are
// the path's own pixel bounds.  The walk starts at the path's top row
// (JobPixBounds) and clips everything else to these.
__host__ __device__ inline PixBounds JobScanBounds(const JobGeo& g) {
  float l = g.l > g.cl ? g.l : g.cl, r = g.r < g.cr ? g.r : g.cr;
  float t = g.t > g.ct ? g.t : g.ct, b = g.b < g.cb ? g.b : g.cb;
  if (!(l < r && t < b)) {
    PixBounds e{0, 0, 0, 0, true};
    return e;
  }
  JobGeo h = g;
  h.l = l;
  h.t = t;
  h.r = r;
  h.b = b;
  return JobPixBounds(h);
}
// Visible tile range of a job on its canvas (inclusive); ntc = 0 if none.
struct TileRange {
  int tc0, tr0, ntc, ntr;
};
__host__ __device__ inline TileRange JobTiles(const PixBounds& pb, int cw, int ch) {
  TileRange tr{0, 0, 0, 0};
  if (pb.empty) return tr;
  int x0 = pb.l > 0 ? pb.l : 0, x1 = pb.r < cw ? pb.r : cw;
  int y0 = pb.t > 0 ? pb.t : 0, y1 = pb.b < ch ? pb.b : ch;
  if (x0 >= x1 || y0 >= y1) return tr;
  tr.tc0 = x0 >> 4;
  tr.tr0 = y0 >> 4;
  tr.ntc = ((x1 - 1) >> 4) - tr.tc0 + 1;
  tr.ntr = ((y1 - 1) >> 4) - tr.tr0 + 1;
  return tr;
}

// Number of row bands of a job and the nominal first row of band b >= 1,
// rounded up to a tile row.
__host__ __device__ inline int BandCount(const PixBounds& pb, int band_rows) {
  if (pb.empty || band_rows <= 0) return 1;
  int nb = (pb.b - pb.t) / band_rows;
  return nb < 1 ? 1 : (nb > kMaxBands ? kMaxBands : nb);
}
__host__ __device__ inline int BandStartRow(const PixBounds& pb, int nb, int b) {
  int v = pb.t + (int)((int64_t)(pb.b - pb.t) * b / nb);
  return (v + kTile - 1) & ~(kTile - 1);
}

// meta word of an active edge: bit 0 winding < 0, bit 1 quad, bits 2.. index
__device__ __forceinline__ int MetaWinding(int m) { return (m & 1) ? -1 : 1; }
__device__ __forceinline__ bool MetaQuad(int m) { return (m & 2) != 0; }
__device__ __forceinline__ int MetaIndex(int m) { return (int)((uint32_t)m >> 2); }
__device__ __forceinline__ int MakeMeta(int winding, bool quad, int idx) {
  return (winding < 0 ? 1 : 0) | (quad ? 2 : 0) | (idx << 2);
}

struct WalkLds {
  int32_t* d;      // packed row deltas (RowSink)
  int32_t* slot0;  // per tile column of the current band
  int32_t* slot1;
  int32_t* rows;   // per tile column: rows written, plane 0 (bits 0-15), plane 1 (16-31)
  int32_t* sX;     // permutation scratch (CAP entries each)
  int32_t* sDX;
  int32_t* sDY;
  int32_t* sLY;
  int32_t* sM;
};

// Bitonic sort (ascending) of N = 64 * 2^j keys held as element
// p = slot * 64 + lane; keys at p >= N are not touched (N <= 64 * K).
template <int K, int SS>
__device__ __forceinline__ void SlotExchange(int64_t (&key)[K], int size, int lane) {
#pragma unroll
  for (int s = 0; s < K; s++) {
    const int t = s ^ SS;
    if (t > s && t < K) {
      int p = s * kWave + lane;
      bool asc = (p & size) == 0;
      int64_t a = key[s], b = key[t];
      bool sw = asc ? (a > b) : (a < b);
      key[s] = sw ? b : a;
      key[t] = sw ? a : b;
    }
  }
}
template <int K>
__device__ void BitonicSortSlots(int64_t (&key)[K], int N, int lane) {
  for (int size = 2; size <= N; size <<= 1) {
    for (int stride = size >> 1; stride > 0; stride >>= 1) {
      if (stride >= kWave) {
        int ss = stride / kWave;
        if (ss == 1) SlotExchange<K, 1>(key, size, lane);
        else if (ss == 2) SlotExchange<K, 2>(key, size, lane);
        else if (ss == 4) SlotExchange<K, 4>(key, size, lane);
        else if (ss == 8) SlotExchange<K, 8>(key, size, lane);
        else if (ss == 16) SlotExchange<K, 16>(key, size, lane);
        else if (ss == 32) SlotExchange<K, 32>(key, size, lane);
      } else {
        bool lower = (lane & stride) == 0;
#pragma unroll
        for (int s = 0; s < K; s++) {
          if (s * kWave >= N) break;
          int64_t o = __shfl_xor(key[s], stride, kWave);
          int p = s * kWave + lane;
          bool asc = (p & size) == 0;
          int64_t mn = key[s] < o ? key[s] : o, mx = key[s] < o ? o : key[s];
          key[s] = (lower == asc) ? mn : mx;
        }
      }
    }
  }
}

// Write pass (kWalkWrite): one wave per (job, band) item; the job's map
// starts at -1 (filled before the launch), its JobCov at zero, and planes are
// taken from the job's counter (JobCov::planes) with atomics, so the bands of
// a job run concurrently.  Count and schedule passes: one wave per job, band
// 0, walking every row and recording the band snapshots.
template <int K, int kMode>
__device__ __forceinline__ void WalkItem(const WalkParams& P, int wi, int32_t* lds) {
  constexpr int CAP = K * kWave;
  constexpr bool kWrite = kMode == kWalkWrite;
  constexpr bool kBlit = kMode != kWalkSchedule;
  const int lane = Lane();
  const uint64_t t_start = wall_clock64();
  const int item = P.walk_list[wi];
  const int j = item >> kBandBits;
  const int rband = kWrite ? (item & (kMaxBands - 1)) : 0;
  const Job job = P.jobs[j];
  const JobGeo geo = P.geo[j];
  const Canvas cv = P.canvases[job.canvas];
  const int path_top = JobPixBounds(geo).t;  // RastePath's start_y
  PixBounds pb = JobScanBounds(geo);
  TileRange tr = JobTiles(pb, cv.width, cv.height);
  // The LDS row holds the visible columns [ox, ox + W) of the canvas
  // (RowSink folds the deltas of off-canvas pixels), so its width is bounded
  // by the canvas, not by the path.
  const int ox = pb.l > 0 ? pb.l : 0;
  const int W = (pb.r < cv.width ? pb.r : cv.width) - ox;
  int flags = 0;
  int n_tiles = 0, max_active = 0;
  int nE = geo.edges;
  int32_t* map0 = kWrite ? P.map0 + P.map_base[j] : nullptr;
  int32_t* map1 = kWrite ? P.map1 + P.map_base[j] : nullptr;
  if (pb.empty || tr.ntc == 0 || nE == 0 || W > P.lds_width) {
    if (!pb.empty && W > P.lds_width) flags |= 16;
    if (lane == 0 && rband == 0) {
      // field by field: a struct copy would go through scratch memory
      JobCov* c = &P.cov[j];
      c->left = pb.l;
      c->top = pb.t;
      c->right = pb.r;
      c->bottom = pb.b;
      c->tiles = 0;
      c->planes = 0;
      c->max_active = 0;
      c->flags = flags;
      c->steps = 0;
      c->time_10ns = 0;
#pragma unroll
      for (int k = 0; k < 5; k++) c->phase[k] = 0;
    }
    return;
  }
  const int nb = P.band_count[j];
  uint8_t* sbase = nb > 1 ? P.snaps + P.snap_base[j] : nullptr;
  auto snap_at = [&](int b) -> SnapHeader* {
    return (SnapHeader*)(sbase + (int64_t)(b - 1) * SnapStride(CAP));
  };
  // the band ends where the next band's snapshot was taken
  const Fixed stop_y = (kWrite && rband + 1 < nb) ? snap_at(rband + 1)->y : kMaxS32;
  WalkLds L;
  L.d = lds;
  L.slot0 = L.d + (P.lds_width + 1);
  L.slot1 = L.slot0 + (P.lds_width / kTile + 2);
  L.rows = L.slot1 + (P.lds_width / kTile + 2);
  L.sX = L.rows + (P.lds_width / kTile + 2);
  L.sDX = L.sX + CAP;
  L.sDY = L.sDX + CAP;
  L.sLY = L.sDY + CAP;
  L.sM = L.sLY + CAP;
  for (int i = lane; i <= W; i += kWave) {
    L.d[i] = 0;
  }
  for (int i = lane; i < tr.ntc; i += kWave) {
    L.slot0[i] = -1;
    L.slot1[i] = -1;
    L.rows[i] = 0;
  }
  WaveSync();

  const EdgeRec* E = P.wedges + P.edge_base[j];
  EdgeRec* Q = P.qstate + P.q_base[j] + (int64_t)rband * nE;
  const int64_t pbase = kWrite ? P.plane_base[j] : 0;
  const int pcap = kWrite ? P.plane_cap[j] : 0;
  uint8_t* planes = P.planes;

  const Fixed left_clip = (Fixed)((uint32_t)pb.l << 16);
  const Fixed right_clip = (Fixed)((uint32_t)pb.r << 16);
  const int winding_mask = job.fill == HIPR_FILL_EVEN_ODD ? 1 : -1;
  const int cx0 = pb.l > 0 ? pb.l : 0, cx1 = pb.r < cv.width ? pb.r : cv.width;
  const int cy0 = pb.t > 0 ? pb.t : 0, cy1 = pb.b < cv.height ? pb.b : cv.height;
  // A job whose blend mode is not a no-op for a transparent source keeps the
  // reference's zero-cover direct spans: its tiles take three planes (direct
  // covers, accumulated covers, and a byte per pixel marking a direct span).
  const bool triple = (job.flags & kJobKeepZero) != 0;
  RowSink sink{L.d, ox, W, triple};

  // Active list registers.
  Fixed X[K], DX[K], DY[K], LY[K];
  int M[K];
  int n = 0;
  int next_slot = 0;  // tile planes allocated so far
  int band = -1;      // current 16-row band
  int dmin = 0x7FFFFFFF, dmax = -0x7FFFFFFF;  // dirty x range of this row

  // ---- helpers ---------------------------------------------------------------
  // Write the active list (m elements in lane/slot layout) from scratch.
  auto load_from_scratch = [&](int m) {
#pragma unroll
    for (int s = 0; s < K; s++) {
      int p = s * kWave + lane;
      if (p < m) {
        X[s] = L.sX[p];
        DX[s] = L.sDX[p];
        DY[s] = L.sDY[p];
        LY[s] = L.sLY[p];
        M[s] = L.sM[p];
      }
    }
  };

  // Finish a band: publish the tile slots of every column to the job map.
  // Each plane row is written once: with coverage by the row flush, or with
  // zeros here for the rows the band never wrote.
  auto finish_band = [&]() {
    if (band < 0) return;
    int trow = band - tr.tr0;
    for (int c = lane; c < tr.ntc; c += kWave) {
      int s0 = L.slot0[c], s1 = L.slot1[c];
      // Only planes that were found are published; a job that found none is
      // reported through flags and redone by the exact path.
      if (kWrite && s0 >= 0 && s0 != kNoPlane) {
        map0[(int64_t)trow * tr.ntc + c] = s0;
        map1[(int64_t)trow * tr.ntc + c] = (s1 >= 0 && s1 != kNoPlane) ? s1 : -1;
      }
      if (kWrite) {
        const uint32_t written = (uint32_t)L.rows[c];
        const int4 zero = {0, 0, 0, 0};
        if (s0 >= 0 && s0 != kNoPlane) {
          for (int t = 0; t < (triple ? 3 : 1); t++) {
            int4* pl = (int4*)(planes + (int64_t)(s0 + t) * kTilePlaneBytes);
            for (int k = 0; k < kTile; k++)
              if (!((written >> k) & 1)) pl[k] = zero;
          }
        }
        if (s1 >= 0 && s1 != kNoPlane) {
          int4* pl = (int4*)(planes + (int64_t)s1 * kTilePlaneBytes);
          for (int k = 0; k < kTile; k++)
            if (!((written >> (16 + k)) & 1)) pl[k] = zero;
        }
      }
      L.slot0[c] = -1;
      L.slot1[c] = -1;
      L.rows[c] = 0;
    }
    WaveSync();
  };

  // Resolve pixel row r (prefix-sum the deltas) and emit its tile rows.
  auto flush_row = [&](int r) {
    int lo = WaveMin(dmin), hi = WaveMax(dmax);
    dmin = 0x7FFFFFFF;
    dmax = -0x7FFFFFFF;
    if (lo > hi) return;  // nothing drawn in this row
    WaveSync();
    bool visible = r >= cy0 && r < cy1;
    int rb = r >> 4;
    if (visible && rb != band) {
      finish_band();
      band = rb;
    }
    // canvas-aligned 16-pixel chunks; lane owns 4 pixels
    int base = (ox + lo) & ~15;
    int end = ox + hi;  // exclusive pixel bound of the dirty range
    int carry = 0;
    // deltas left of `base` are zero: base <= ox + lo
    for (int x0 = base; x0 <= end; x0 += 4 * kWave) {
      int px = x0 + 4 * lane;
      int32_t vS[4];
      int sS = 0;
#pragma unroll
      for (int q = 0; q < 4; q++) {
        int i = px + q - ox;
        int dv = 0;
        if (i >= 0 && i <= W) {
          dv = L.d[i];
          L.d[i] = 0;
        }
        sS += dv;
        vS[q] = sS;
      }
      int iS = ScanSum(sS);
      int eS = iS - sS + carry;
      carry += __builtin_amdgcn_readlane(iS, kWave - 1);
      uint32_t c0 = 0, c1 = 0, c2 = 0;
      bool stacked = false;
#pragma unroll
      for (int q = 0; q < 4; q++) {
        int S = vS[q] + eS;
        int a = S & 0xFFFF, d = S >> 16;
        int cnt = d >> 12, dv = d & 0xFFF;
        int av = a > 255 ? 255 : a;
        int first = cnt ? dv : av;
        int second = (cnt && av) ? av : 0;
        int third = 0;
        if (triple) {
          first = cnt ? dv : 0;
          second = av;
          third = cnt ? 1 : 0;
        }
        if (cnt >= 2) stacked = true;
        int x = px + q;
        if (!visible || x < cx0 || x >= cx1) {
          first = 0;
          second = 0;
          third = 0;
        }
        c0 |= (uint32_t)(first & 0xFF) << (8 * q);
        c1 |= (uint32_t)(second & 0xFF) << (8 * q);
        c2 |= (uint32_t)third << (8 * q);
      }
      if (stacked) flags |= 4;
      if (!visible) continue;
      // chunk (4 lanes = 16 px) status
      uint64_t nz0 = __ballot(triple ? (c0 | c1 | c2) != 0 : c0 != 0);
      uint64_t nz1 = __ballot(!triple && c1 != 0);
      int g = lane >> 2;
      bool chunk0 = ((nz0 >> (4 * g)) & 0xF) != 0;
      bool chunk1 = ((nz1 >> (4 * g)) & 0xF) != 0;
      int tcol = (px >> 4) - tr.tc0;
      bool in_cols = tcol >= 0 && tcol < tr.ntc;
      bool leader = (lane & 3) == 0;
      // allocate planes for chunks that become non-empty in this band
      int cur0 = in_cols ? L.slot0[tcol] : -1;
      int cur1 = in_cols ? L.slot1[tcol] : -1;
      bool need0 = in_cols && leader && chunk0 && cur0 < 0;
      bool need1 = in_cols && leader && chunk1 && cur1 < 0;
      // most chunks already hold their planes: allocate only when needed
      if (__ballot(need0 || need1) != 0) {
        int req = triple ? (need0 ? 3 : 0) : (need0 ? 1 : 0) + (need1 ? 1 : 0);
        int rincl = ScanSum(req);
        int rtot = __builtin_amdgcn_readlane(rincl, kWave - 1);
        int slot_base = next_slot;
        if (kWrite && rtot > 0) {
          // the bands of a job share its plane range
          int got = 0;
          if (lane == 0) got = atomicAdd(&P.cov[j].planes, rtot);
          slot_base = __builtin_amdgcn_readlane(got, 0);
        }
        int first_slot = slot_base + rincl - req;
        next_slot += rtot;
        // Count pass: tile slots are job-relative.  Write pass: absolute plane
        // numbers, from the job's range or (past its capacity) from the pool.
        int a0 = first_slot, a1 = first_slot + (need0 ? 1 : 0);
        if (kWrite) {
          const int over_from = slot_base > pcap ? slot_base : pcap;
          const int over = slot_base + rtot - over_from;
          int pool0 = 0;
          if (over > 0) {
            int got = 0;
            if (lane == 0) got = atomicAdd(P.pool_ctr, over);
            pool0 = __builtin_amdgcn_readlane(got, 0);
          }
          auto to_abs = [&](int sl) -> int {
            if (sl < pcap) return (int)(pbase + sl);
            int pi = pool0 + (sl - over_from);
            return pi < P.pool_cap ? P.pool_base + pi : kNoPlane;
          };
          // a triple must be contiguous: it never comes from the pool (a
          // job that outgrows its range that way is walked again exactly)
          a0 = triple ? (a0 + 2 < pcap ? (int)(pbase + a0) : kNoPlane) : to_abs(a0);
          a1 = to_abs(a1);
        }
        if (need0) {
          L.slot0[tcol] = a0;
          n_tiles++;
        }
        if (need1) L.slot1[tcol] = a1;
        WaveSync();
        cur0 = in_cols ? L.slot0[tcol] : -1;
        cur1 = in_cols ? L.slot1[tcol] : -1;
      }
      if (kWrite) {
        int sub = lane & 3;
        int rowk = r & 15;
        if (in_cols) {
          if (sub == 0) {
            // rows written in this band (the other rows are zeroed when the
            // band ends); one lane per tile column
            int wr = (chunk0 && cur0 >= 0 ? 1 << rowk : 0) | (chunk1 && cur1 >= 0 ? 1 << (16 + rowk) : 0);
            if (wr) L.rows[tcol] |= wr;
          }
          if (chunk0 && cur0 >= 0) {
            if (cur0 != kNoPlane) {
              uint32_t* pl = (uint32_t*)(planes + (int64_t)cur0 * kTilePlaneBytes);
              pl[rowk * 4 + sub] = c0;
              if (triple) {
                pl[kTilePlaneBytes / 4 + rowk * 4 + sub] = c1;
                pl[2 * kTilePlaneBytes / 4 + rowk * 4 + sub] = c2;
              }
            } else {
              flags |= 1;
            }
          }
          if (chunk1 && cur1 >= 0) {
            if (cur1 != kNoPlane) {
              uint32_t* pl = (uint32_t*)(planes + (int64_t)cur1 * kTilePlaneBytes);
              pl[rowk * 4 + sub] = c1;
            } else {
              flags |= 1;
            }
          }
        }
      }
    }
    // clear the end-of-row delta entry
    if (lane == 0) L.d[W] = 0;
    WaveSync();
  };

  // Insert `k` new edges already in scratch (sX.. at [CAP-k.. ) no: at 0..k-1
  // of the n* arrays) -- implemented inline below.

  // ---- pending edges ---------------------------------------------------------
  int pp = 0;  // next pending edge (sorted position)
  int pu_idx = -1;
  Fixed pu_val = kMaxS32;  // upper_y of edge pu_idx, read once
  auto pending_upper = [&](int idx) -> Fixed {
    if (idx != pu_idx) {
      pu_idx = idx;
      pu_val = idx < nE ? E[idx].upper_y : kMaxS32;
    }
    return pu_val;
  };

  // Gathers the new edges with upper_y <= y into scratch slots
  // [CAP/2 ...] -- uses separate regions of the scratch arrays: survivors are
  // reloaded before new edges are staged, so the whole scratch is free.
  // Returns the number of new edges (sorted by x, dx, index).
  auto gather_new = [&](Fixed y, int* out_k) -> bool {
    int k = 0;
    for (;;) {
      int idx = pp + k + lane;
      Fixed uy = idx < nE ? E[idx].upper_y : kMaxS32;
      bool ready = uy <= y;
      uint64_t b = __ballot(ready);
      int cnt = __popcll(b);
      if (k + cnt > CAP) return false;
      if (ready) {
        const EdgeRec* er = &E[idx];
        const int32_t packed = er->packed;
        if (EdgeIsQuad(packed)) {
          // the quad's state from here on; four 16-byte moves (a struct
          // copy would be staged through scratch memory)
          const int4* src = (const int4*)er;
          int4* dst = (int4*)&Q[idx];
          const int4 q0 = src[0], q1 = src[1], q2 = src[2], q3 = src[3];
          dst[0] = q0;
          dst[1] = q1;
          dst[2] = q2;
          dst[3] = q3;
        }
        int p = k + lane;
        L.sX[p] = er->x;
        L.sDX[p] = er->dx;
        L.sDY[p] = er->dy;
        L.sLY[p] = er->lower_y;
        L.sM[p] = MakeMeta(EdgeWinding(packed), EdgeIsQuad(packed), idx);
      }
      k += cnt;
      if (cnt < kWave) break;
    }
    *out_k = k;
    return true;
  };

  // ---- initial activation (sw_raster.cc:554-563) -------------------------------
  Fixed y = Max32(E[0].upper_y, IntToFixed(path_top));
  Fixed nny = kMaxS32;
  if (rband > 0) {
    // resume from the band's snapshot
    const SnapHeader* h = snap_at(rband);
    y = h->y;
    if (y == kMaxS32) return;  // the walk ended before this band
    nny = h->nny;
    pp = h->pp;
    n = h->n;
    const int32_t* a = (const int32_t*)(h + 1);
    const EdgeRec* qs = (const EdgeRec*)(a + 5 * CAP);
#pragma unroll
    for (int s = 0; s < K; s++) {
      int p = s * kWave + lane;
      if (p < n) {
        X[s] = a[p];
        DX[s] = a[CAP + p];
        DY[s] = a[2 * CAP + p];
        LY[s] = a[3 * CAP + p];
        M[s] = a[4 * CAP + p];
        if (MetaQuad(M[s])) Q[MetaIndex(M[s])] = qs[p];
      }
    }
  } else {
    int k;
    if (!gather_new(y, &k)) {
      flags |= 2;
      k = 0;
    }
    WaveSync();
    // sort the k new edges by (x, dx, index): rank by counting
    // (the sorted-by-upper_y input keeps index order within equal keys)
    int rk[K];
#pragma unroll
    for (int s = 0; s < K; s++) {
      int p = s * kWave + lane;
      rk[s] = 0;
      if (p < k) {
        Fixed xi = L.sX[p], dxi = L.sDX[p];
        int r = 0;
        for (int q = 0; q < k; q++) {
          Fixed xq = L.sX[q], dxq = L.sDX[q];
          bool less = xq < xi || (xq == xi && (dxq < dxi || (dxq == dxi && q < p)));
          r += less ? 1 : 0;
        }
        rk[s] = r;
      }
    }
    Fixed tX[K], tDX[K], tDY[K], tLY[K];
    int tM[K];
#pragma unroll
    for (int s = 0; s < K; s++) {
      int p = s * kWave + lane;
      if (p < k) {
        tX[s] = L.sX[p];
        tDX[s] = L.sDX[p];
        tDY[s] = L.sDY[p];
        tLY[s] = L.sLY[p];
        tM[s] = L.sM[p];
      }
    }
    WaveSync();
#pragma unroll
    for (int s = 0; s < K; s++) {
      int p = s * kWave + lane;
      if (p < k) {
        L.sX[rk[s]] = tX[s];
        L.sDX[rk[s]] = tDX[s];
        L.sDY[rk[s]] = tDY[s];
        L.sLY[rk[s]] = tLY[s];
        L.sM[rk[s]] = tM[s];
      }
    }
    WaveSync();
    load_from_scratch(k);
    n = k;
    pp = k;
    int mn = kMaxS32;
#pragma unroll
    for (int s = 0; s < K; s++) {
      int p = s * kWave + lane;
      if (p < n && LY[s] > y) mn = Min32(mn, LY[s]);
    }
    nny = WaveMin(mn);
    Fixed nu = pending_upper(pp);
    if (nu > y && nu < nny) nny = nu;
    WaveSync();
  }

  const Fixed stop_fixed = IntToFixed(pb.b);
  int guard = 0;
  uint64_t ph[5] = {0, 0, 0, 0, 0};  // phase times (10 ns units)
  uint64_t tp_loop = wall_clock64();
  int next_b = 1;  // next band snapshot to record (count / schedule passes)
  for (;;) {
    if (y >= stop_y) break;  // the next band starts here
    if (!kWrite && next_b < nb && (y & 0xFFFF) == 0 && ((y >> 16) & (kTile - 1)) == 0) {
      const int yrow = y >> 16;
      while (next_b < nb && yrow >= BandStartRow(pb, nb, next_b)) {
        SnapHeader* h = snap_at(next_b);
        if (lane == 0) {
          h->y = y;
          h->nny = nny;
          h->pp = pp;
          h->n = n;
        }
        int32_t* a = (int32_t*)(h + 1);
        EdgeRec* qs = (EdgeRec*)(a + 5 * CAP);
#pragma unroll
        for (int s = 0; s < K; s++) {
          int p = s * kWave + lane;
          if (p < n) {
            a[p] = X[s];
            a[CAP + p] = DX[s];
            a[2 * CAP + p] = DY[s];
            a[3 * CAP + p] = LY[s];
            a[4 * CAP + p] = M[s];
            if (MetaQuad(M[s])) qs[p] = Q[MetaIndex(M[s])];
          }
        }
        next_b++;
      }
    }
    if (++guard > (1 << 26)) {  // cannot happen: y strictly increases
      flags |= 32;
      break;
    }
    max_active = Max32(max_active, n);
    Fixed next_y = Min32(nny, FixedCeilToFixed(y + 1));
    int y_shift = 0;
    if ((next_y - y) & (kFixed1 >> 2)) {
      y_shift = 2;
      next_y = y + (kFixed1 >> 2);
    } else if ((next_y - y) & (kFixed1 >> 1)) {
      y_shift = 1;
    }
    const Alpha full = (Alpha)FixedRoundToInt(0xFF * (next_y - y));
    const int row = y >> 16;
    const int S = (n + kWave - 1) / kWave;

    uint64_t tp0 = wall_clock64();
    Fixed XN[K];
    if (!kBlit) {
#pragma unroll
      for (int s = 0; s < K; s++) XN[s] = X[s] + (DX[s] >> y_shift);
    }
    // ---- phases 1-3: roles, trapezoids --------------------------------------
    if (kBlit) {
    int carry_w = 0;
    int carry_in = 0;
    int carry_lidx = -1;
    Fixed carry_lxo = 0, carry_lxn = 0, carry_ldy = 0;
    int carry_ridx = -1;
    Fixed carry_rxo = 0, carry_rxn = 0;
#pragma unroll
    for (int s = 0; s < K; s++) {
      if (s >= S) break;
      int p = s * kWave + lane;
      bool valid = p < n;
      int wv = valid ? MetaWinding(M[s]) : 0;
      int incl = ScanSum(wv) + carry_w;
      int in = (valid && (incl & winding_mask) != 0) ? 1 : 0;
      int prev_in = ShiftUp(in, carry_in);
      bool is_left = valid && in && !prev_in;
      bool is_right = valid && !in && prev_in;
      Fixed xo = X[s];
      Fixed xn = valid ? xo + (DX[s] >> y_shift) : 0;
      XN[s] = xn;
      // nearest left edge at or before p
      int lsel = ScanMax(is_left ? lane : -1);
      Fixed lxo = __shfl(xo, lsel < 0 ? 0 : lsel, kWave);
      Fixed lxn = __shfl(xn, lsel < 0 ? 0 : lsel, kWave);
      Fixed ldy = __shfl(DY[s], lsel < 0 ? 0 : lsel, kWave);
      if (lsel < 0) {
        lxo = carry_lxo;
        lxn = carry_lxn;
        ldy = carry_ldy;
      }
      // previous right edge strictly before p
      int rsel_incl = ScanMax(is_right ? lane : -1);
      int rsel = ShiftUp(rsel_incl, -1);
      Fixed rxo = __shfl(xo, rsel < 0 ? 0 : rsel, kWave);
      Fixed rxn = __shfl(xn, rsel < 0 ? 0 : rsel, kWave);
      bool has_prev_right = rsel >= 0 || carry_ridx >= 0;
      if (rsel < 0) {
        rxo = carry_rxo;
        rxn = carry_rxn;
      }
      // next element (old x, dx); lane 63's successor is lane 0 of slot s+1
      Fixed nxo = ShiftDown(xo, xo), ndx = ShiftDown(DX[s], DX[s]);
      int next_p = p + 1;
      if (s + 1 < K) {
        Fixed n0x = __builtin_amdgcn_readlane(X[s + 1 < K ? s + 1 : s], 0);
        Fixed n0dx = __builtin_amdgcn_readlane(DX[s + 1 < K ? s + 1 : s], 0);
        if (lane == kWave - 1) {
          nxo = n0x;
          ndx = n0dx;
        }
      }
      bool next_active = next_p < n;
      if (is_right) {
        Fixed left = Max32(lxo, left_clip);
        Fixed right = Min32(right_clip, xo);
        Fixed next_left = Max32(left_clip, lxn);
        Fixed next_right = Min32(right_clip, xn);
        int prev_right = has_prev_right
                             ? FixedCeilToInt(Max32(Min32(right_clip, rxo), rxn))
                             : FixedFloorToInt(left_clip);
        bool too_close_prev =
            prev_right > FixedFloorToInt(left) || prev_right > FixedFloorToInt(lxn);
        bool too_close_next = next_active && (xn + kFixed1 >= nxo - Abs32(ndx));
        bool no_real = full == 0xFF && (too_close_prev || too_close_next);
        BlitTrapezoidRow(&sink, left, right, next_left, next_right, ldy, DY[s], full,
                         no_real);
        int lo = FixedFloorToInt(Min32(Min32(left, next_left), Min32(right, next_right))) - ox;
        int hi = FixedCeilToInt(Max32(Max32(left, next_left), Max32(right, next_right))) - ox;
        dmin = Min32(dmin, Max32(lo, 0));
        dmax = Max32(dmax, Min32(hi, W));
      }
      // carries
      carry_w = __builtin_amdgcn_readlane(incl, kWave - 1);
      carry_in = __builtin_amdgcn_readlane(in, kWave - 1);
      int lastl = __builtin_amdgcn_readlane(lsel, kWave - 1);
      if (lastl >= 0) {
        carry_lidx = s * kWave + lastl;
        carry_lxo = __builtin_amdgcn_readlane(xo, lastl);
        carry_lxn = __builtin_amdgcn_readlane(xn, lastl);
        carry_ldy = __builtin_amdgcn_readlane(DY[s], lastl);
      }
      int lastr = __builtin_amdgcn_readlane(rsel_incl, kWave - 1);
      if (lastr >= 0) {
        carry_ridx = s * kWave + lastr;
        carry_rxo = __builtin_amdgcn_readlane(xo, lastr);
        carry_rxn = __builtin_amdgcn_readlane(xn, lastr);
      }
    }
    // open interval at the end of the list (right edge culled away)
    if (carry_in) {
      flags |= 8;
      // The reference tests edges_too_close(left_edge->prev, left_edge);
      // with a closed path the winding sum is 0 here, so this is not reached
      // on the parity scenes.  The left edge's list predecessor is not
      // tracked; the span is emitted as accumulated coverage.
      Fixed left = Max32(carry_lxo, left_clip);
      if (lane == 0) {
        BlitTrapezoidRow(&sink, left, right_clip, Max32(left_clip, carry_lxn), right_clip,
                         carry_ldy, 0, full, full == 0xFF);
        int lo = FixedFloorToInt(Min32(left, Max32(left_clip, carry_lxn))) - ox;
        dmin = Min32(dmin, Max32(lo, 0));
        dmax = Max32(dmax, W);
      }
    }
    WaveSync();
    }  // kBlit

    uint64_t tp1 = wall_clock64();
    ph[0] += tp1 - tp0;
    // ---- phase 4: GoY and quad updates -----------------------------------------
#pragma unroll
    for (int s = 0; s < K; s++) {
      if (s >= S) break;
      int p = s * kWave + lane;
      if (p < n) {
        X[s] = XN[s];
        if (LY[s] <= next_y && MetaQuad(M[s])) {
          // while (lower_y <= next_y) { KeepContinuous; UpdateQuad } with
          // KeepContinuous taking the edge's current x and y
          // (sw_raster.cc:628-638, sw_edge.cc:294-297).
          int idx = MetaIndex(M[s]);
          EdgeRec e = Q[idx];
          Fixed sy = next_y;
          Fixed cx = X[s];
          while (LY[s] <= next_y && EdgeCurveCount(e.packed) > 0) {
            e.snapped_x = cx;
            if (!UpdateQuad(&e, sy)) break;
            cx = e.x;
            sy = e.upper_y;
            DX[s] = e.dx;
            DY[s] = e.dy;
            LY[s] = e.lower_y;
          }
          X[s] = cx;
          M[s] = MakeMeta(EdgeWinding(e.packed), true, idx);
          Q[idx] = e;
        }
      }
    }

    uint64_t tp2 = wall_clock64();
    ph[1] += tp2 - tp1;
    // ---- phases 5-6: survivors, stable x order, intersection checks ----------
    int m = 0;
    bool flag_isect = false;
    {
      int cnt_carry = 0;
      bool sorted = true;
      Fixed prevX_carry = kMinS32;
      bool has_prev_carry = false;
      int C[K];
      bool SV[K];
#pragma unroll
      for (int s = 0; s < K; s++) {
        if (s >= S) break;
        int p = s * kWave + lane;
        bool sv = p < n && LY[s] > next_y;
        SV[s] = sv;
        int inc = ScanSum(sv ? 1 : 0);
        C[s] = cnt_carry + inc - (sv ? 1 : 0);
        // previous survivor's x
        int psel_incl = ScanMax(sv ? lane : -1);
        int psel = ShiftUp(psel_incl, -1);
        Fixed px = __shfl(X[s], psel < 0 ? 0 : psel, kWave);
        Fixed pdx = __shfl(DX[s], psel < 0 ? 0 : psel, kWave);
        bool hp = psel >= 0;
        if (!hp) {
          px = prevX_carry;
          hp = has_prev_carry;
        }
        if (sv && hp && X[s] < px) sorted = false;
        (void)pdx;
        int lastp = __builtin_amdgcn_readlane(psel_incl, kWave - 1);
        if (lastp >= 0) {
          prevX_carry = __builtin_amdgcn_readlane(X[s], lastp);
          has_prev_carry = true;
        }
        cnt_carry += __builtin_amdgcn_readlane(inc, kWave - 1);
      }
      m = cnt_carry;
      bool all_sorted = __ballot(!sorted) == 0;
      if (all_sorted && m == n) {
        // every edge survives in place: the predecessor of element p is
        // p - 1 (check_intersection against the previous element)
        bool f = false;
        Fixed carry_v = 0;
#pragma unroll
        for (int s = 0; s < K; s++) {
          if (s >= S) break;
          int p = s * kWave + lane;
          Fixed v = X[s] + DX[s];
          Fixed pv = ShiftUp(v, carry_v);
          if (p >= 1 && p < n && pv > v) f = true;
          carry_v = __builtin_amdgcn_readlane(v, kWave - 1);
        }
        flag_isect = __ballot(f) != 0;
      } else {
      // write survivors to scratch in survivor order
#pragma unroll
      for (int s = 0; s < K; s++) {
        if (s >= S) break;
        if (SV[s]) {
          int c = C[s];
          L.sX[c] = X[s];
          L.sDX[c] = DX[s];
          L.sDY[c] = DY[s];
          L.sLY[c] = LY[s];
          L.sM[c] = M[s];
        }
      }
      WaveSync();
      if (all_sorted) {
        // predecessor = previous survivor
        bool f = false;
        for (int c = lane + 1; c < m; c += kWave) {
          if (L.sX[c - 1] + L.sDX[c - 1] > L.sX[c] + L.sDX[c]) f = true;
        }
        flag_isect = __ballot(f) != 0;
        load_from_scratch(m);
      } else {
        // Stable re-sort by x.  The survivors stay in scratch in list order;
        // only the keys (x, list position) are sorted in registers, by
        // odd-even transposition (a step only swaps edges that crossed, so
        // one or two passes usually do), with a bitonic network when eight
        // passes do not; then the other fields are gathered by position.
        int O[K];
#pragma unroll
        for (int s = 0; s < K; s++) {
          int p = s * kWave + lane;
          O[s] = p;
          if (p < m) X[s] = L.sX[p];
        }
        const int S2 = (m + kWave - 1) / kWave;
        auto gt_key = [](Fixed xa, int oa, Fixed xb, int ob) {
          return xa > xb || (xa == xb && oa > ob);
        };
        bool sorted = false;
        for (int pass = 0; pass < 8 && !sorted; pass++) {
          // even phase: pairs (2t, 2t+1) are lanes (2t, 2t+1) of one slot
#pragma unroll
          for (int s = 0; s < K; s++) {
            if (s >= S2) break;
            int p = s * kWave + lane;
            int q = p ^ 1;
            Fixed ox = __shfl_xor(X[s], 1, kWave);
            int oo = __shfl_xor(O[s], 1, kWave);
            bool both = p < m && q < m;
            bool left = (lane & 1) == 0;
            bool take = both && (left ? gt_key(X[s], O[s], ox, oo) : gt_key(ox, oo, X[s], O[s]));
            if (take) {
              X[s] = ox;
              O[s] = oo;
            }
          }
          // odd phase: pairs (2t+1, 2t+2); lane 63 of slot s pairs with
          // lane 0 of slot s+1.  Slot s+1 has not moved yet when slot s is
          // processed; lane 63 of slot s-1 is kept from before it moved.
          Fixed p63x = 0;
          int p63o = 0;
#pragma unroll
          for (int s = 0; s < K; s++) {
            if (s >= S2) break;
            int p = s * kWave + lane;
            bool left = (lane & 1) == 1;  // odd position = left of its pair
            int q = left ? p + 1 : p - 1;
            const Fixed c63x = __builtin_amdgcn_readlane(X[s], kWave - 1);
            const int c63o = __builtin_amdgcn_readlane(O[s], kWave - 1);
            Fixed dnx = ShiftDown(X[s], 0), upx = ShiftUp(X[s], 0);
            int dno = ShiftDown(O[s], 0), upo = ShiftUp(O[s], 0);
            if (s + 1 < K) {
              const int t = s + 1 < K ? s + 1 : s;
              const Fixed n0x = __builtin_amdgcn_readlane(X[t], 0);
              const int n0o = __builtin_amdgcn_readlane(O[t], 0);
              if (lane == kWave - 1) {
                dnx = n0x;
                dno = n0o;
              }
            }
            if (lane == 0) {
              upx = p63x;
              upo = p63o;
            }
            bool both = p < m && q >= 0 && q < m;
            bool take = both && (left ? gt_key(X[s], O[s], dnx, dno) : gt_key(upx, upo, X[s], O[s]));
            if (take) {
              X[s] = left ? dnx : upx;
              O[s] = left ? dno : upo;
            }
            p63x = c63x;
            p63o = c63o;
          }
          // sorted when every element exceeds its predecessor
          bool bad = false;
          Fixed cx = 0;
          int co = 0;
#pragma unroll
          for (int s = 0; s < K; s++) {
            if (s >= S2) break;
            int p = s * kWave + lane;
            Fixed px = ShiftUp(X[s], cx);
            int po = ShiftUp(O[s], co);
            if (p >= 1 && p < m && gt_key(px, po, X[s], O[s])) bad = true;
            cx = __builtin_amdgcn_readlane(X[s], kWave - 1);
            co = __builtin_amdgcn_readlane(O[s], kWave - 1);
          }
          sorted = __ballot(bad) == 0;
        }
        if (!sorted) {
          // bitonic sort of the keys (x, list position)
          int64_t key[K];
#pragma unroll
          for (int s = 0; s < K; s++) {
            int c = s * kWave + lane;
            // signed x in the high word, list position (>= 0) in the low word
            key[s] = c < m ? (int64_t)(((uint64_t)(uint32_t)X[s] << 32) | (uint32_t)O[s])
                           : INT64_MAX;
          }
          int N = kWave;
          while (N < m) N <<= 1;
          BitonicSortSlots<K>(key, N, lane);
#pragma unroll
          for (int s = 0; s < K; s++) {
            int c = s * kWave + lane;
            if (c < m) {
              X[s] = (Fixed)(key[s] >> 32);
              O[s] = (int)(uint32_t)((uint64_t)key[s] & 0xFFFFFFFFull);
            }
          }
        }
        // the other fields, by list position
#pragma unroll
        for (int s = 0; s < K; s++) {
          int c = s * kWave + lane;
          if (c < m) {
            DX[s] = L.sDX[O[s]];
            DY[s] = L.sDY[O[s]];
            LY[s] = L.sLY[O[s]];
            M[s] = L.sM[O[s]];
          }
        }
        WaveSync();
        // check_intersection with the predecessor at insertion time: the
        // nearest earlier element of the sorted list whose list position is
        // smaller (edges that moved in front of it were processed later).
#pragma unroll
        for (int s = 0; s < K; s++) {
          int c = s * kWave + lane;
          if (c < m) {
            L.sX[c] = X[s];
            L.sDX[c] = DX[s];
            L.sLY[c] = O[s];
          }
        }
        WaveSync();
        bool f = false;
#pragma unroll
        for (int s = 0; s < K; s++) {
          int c = s * kWave + lane;
          if (c < m) {
            int q = c - 1;
            while (q >= 0 && L.sLY[q] > O[s]) q--;
            if (q >= 0 && L.sX[q] + L.sDX[q] > X[s] + DX[s]) f = true;
          }
        }
        flag_isect = __ballot(f) != 0;
      }
      WaveSync();
      }  // survivors moved
    }
    n = m;
    {
      int mn = kMaxS32;
#pragma unroll
      for (int s = 0; s < K; s++) {
        int p = s * kWave + lane;
        if (p < n) mn = Min32(mn, LY[s]);
      }
      nny = WaveMin(mn);
      if (flag_isect) nny = next_y + (kFixed1 >> kAcc);
    }

    uint64_t tp3 = wall_clock64();
    ph[2] += tp3 - tp2;
    // ---- row completion ------------------------------------------------------
    y = next_y;
    if (kBlit && (y & 0xFFFF) == 0) flush_row(row);
    uint64_t tp4 = wall_clock64();
    ph[3] += tp4 - tp3;
    if (y >= stop_fixed) break;

    // ---- phase 7: insert new edges (insert_new_edges) --------------------------
    Fixed nu = pending_upper(pp);
    if (nu <= y) {
      int k;
      if (!gather_new(y, &k) || n + k > CAP) {
        flags |= 2;
        break;
      }
      WaveSync();
      // sort the new edges by (x, dx, sorted position)
      int rk[K];
      Fixed tX[K], tDX[K], tDY[K], tLY[K];
      int tM[K];
#pragma unroll
      for (int s = 0; s < K; s++) {
        int p = s * kWave + lane;
        rk[s] = 0;
        if (p < k) {
          Fixed xi = L.sX[p], dxi = L.sDX[p];
          int r = 0;
          for (int q = 0; q < k; q++) {
            Fixed xq = L.sX[q], dxq = L.sDX[q];
            r += (xq < xi || (xq == xi && (dxq < dxi || (dxq == dxi && q < p)))) ? 1 : 0;
          }
          rk[s] = r;
          tX[s] = L.sX[p];
          tDX[s] = L.sDX[p];
          tDY[s] = L.sDY[p];
          tLY[s] = L.sLY[p];
          tM[s] = L.sM[p];
        }
      }
      WaveSync();
#pragma unroll
      for (int s = 0; s < K; s++) {
        int p = s * kWave + lane;
        if (p < k) {
          L.sX[rk[s]] = tX[s];
          L.sDX[rk[s]] = tDX[s];
          L.sDY[rk[s]] = tDY[s];
          L.sLY[rk[s]] = tLY[s];
          L.sM[rk[s]] = tM[s];
        }
      }
      WaveSync();
      // insertion positions: P1 = #{a : x_a <= x_N1};
      // Pi = max(P(i-1), #{a : x_a < x_Ni}).
      // The new edges are at scratch[0..k); positions go to sLY? keep in regs
      int Pcur = 0;
      bool fnew = false;
      int posN[K];
#pragma unroll
      for (int s = 0; s < K; s++) posN[s] = 0;
      for (int i = 0; i < k; i++) {
        Fixed xi = L.sX[i];
        int cnt = 0;
#pragma unroll
        for (int s = 0; s < K; s++) {
          int p = s * kWave + lane;
          bool v = p < n && (i == 0 ? X[s] <= xi : X[s] < xi);
          cnt += __popcll(__ballot(v));
        }
        int Pi = i == 0 ? cnt : Max32(Pcur, cnt);
        // predecessor at insertion: N(i-1) if it sits right before, else A[Pi-1]
        Fixed predx = 0, preddx = 0;
        bool has_pred = false;
        if (i > 0 && Pi == Pcur) {
          predx = L.sX[i - 1];
          preddx = L.sDX[i - 1];
          has_pred = true;
        } else if (Pi > 0) {
          int q = Pi - 1;
          // q is the same for every lane: read the element from its slot
          const int s = __builtin_amdgcn_readfirstlane(q / kWave);
          const int ln = __builtin_amdgcn_readfirstlane(q % kWave);
          Fixed ax = 0, adx = 0;
#pragma unroll
          for (int t = 0; t < K; t++) {
            if (t == s) {
              ax = __builtin_amdgcn_readlane(X[t], ln);
              adx = __builtin_amdgcn_readlane(DX[t], ln);
            }
          }
          predx = ax;
          preddx = adx;
          has_pred = true;
        }
        if (has_pred && predx + preddx > L.sX[i] + L.sDX[i]) fnew = true;
        if (i % kWave == lane) {
#pragma unroll
          for (int s = 0; s < K; s++)
            if (s == i / kWave) posN[s] = Pi + i;
        }
        Pcur = Pi;
      }
      if (fnew) nny = y + (kFixed1 >> kAcc);
      // lower_y of new edges
      {
        int mn = kMaxS32;
#pragma unroll
        for (int s = 0; s < K; s++) {
          int p = s * kWave + lane;
          if (p < k) mn = Min32(mn, L.sLY[p]);
        }
        mn = WaveMin(mn);
        if (mn > y && mn < nny) nny = mn;
      }
      // merge: A[q] moves to q + #{i : P_i <= q}; N_i goes to P_i + i
      // P_i are stored as posN - i; count with a loop over i
      Fixed nX[K], nDX[K], nDY[K], nLY[K];
      int nM[K];
#pragma unroll
      for (int s = 0; s < K; s++) {
        int p = s * kWave + lane;
        if (p < k) {
          nX[s] = L.sX[p];
          nDX[s] = L.sDX[p];
          nDY[s] = L.sDY[p];
          nLY[s] = L.sLY[p];
          nM[s] = L.sM[p];
        }
      }
      // P_i broadcast through LDS (reuse sLY region after reading)
      WaveSync();
#pragma unroll
      for (int s = 0; s < K; s++) {
        int p = s * kWave + lane;
        if (p < k) L.sLY[CAP - 1 - p] = posN[s] - p;  // P_p, stored from the end
      }
      WaveSync();
      int shiftA[K];
#pragma unroll
      for (int s = 0; s < K; s++) {
        int q = s * kWave + lane;
        int c = 0;
        if (q < n) {
          for (int i = 0; i < k; i++) c += (L.sLY[CAP - 1 - i] <= q) ? 1 : 0;
        }
        shiftA[s] = c;
      }
      WaveSync();
#pragma unroll
      for (int s = 0; s < K; s++) {
        int q = s * kWave + lane;
        if (q < n) {
          int d = q + shiftA[s];
          L.sX[d] = X[s];
          L.sDX[d] = DX[s];
          L.sDY[d] = DY[s];
          L.sLY[d] = LY[s];
          L.sM[d] = M[s];
        }
      }
#pragma unroll
      for (int s = 0; s < K; s++) {
        int p = s * kWave + lane;
        if (p < k) {
          int d = posN[s];
          L.sX[d] = nX[s];
          L.sDX[d] = nDX[s];
          L.sDY[d] = nDY[s];
          L.sLY[d] = nLY[s];
          L.sM[d] = nM[s];
        }
      }
      WaveSync();
      n += k;
      pp += k;
      load_from_scratch(n);
      WaveSync();
      Fixed nu2 = pending_upper(pp);
      if (nu2 > y && nu2 < nny) nny = nu2;
    } else if (nu > y && nu < nny) {
      nny = nu;
    }
  }
  finish_band();
  if (!kWrite) {
    // bands the walk never reached are empty
    for (; next_b < nb; next_b++)
      if (lane == 0) snap_at(next_b)->y = kMaxS32;
  }
  n_tiles = WaveSum(n_tiles);
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) flags |= __shfl_xor(flags, o, kWave);
  if (lane == 0) {
    uint64_t t_end = wall_clock64();
    uint64_t dt = t_end - t_start;
    ph[4] = (t_end - tp_loop) - (ph[0] + ph[1] + ph[2] + ph[3]);
    int32_t dt32 = (int32_t)(dt < 0x7FFFFFFF ? dt : 0x7FFFFFFF);
    if (kWrite) {
      JobCov* c = &P.cov[j];
      c->left = pb.l;
      c->top = pb.t;
      c->right = pb.r;
      c->bottom = pb.b;
      atomicAdd(&c->tiles, n_tiles);
      atomicMax(&c->max_active, max_active);
      atomicOr(&c->flags, flags);
      atomicAdd(&c->steps, guard);
      atomicMax(&c->time_10ns, dt32);
#pragma unroll
      for (int k = 0; k < 5; k++)
        atomicAdd(&c->phase[k], (int32_t)(ph[k] < 0x7FFFFFFF ? ph[k] : 0x7FFFFFFF));
    } else {
      JobCov* c = &P.cov[j];
      c->left = pb.l;
      c->top = pb.t;
      c->right = pb.r;
      c->bottom = pb.b;
      c->tiles = n_tiles;
      c->planes = next_slot;
      c->max_active = max_active;
      c->flags = flags;
      c->steps = guard;
      c->time_10ns = dt32;
#pragma unroll
      for (int k = 0; k < 5; k++) c->phase[k] = (int32_t)(ph[k] < 0x7FFFFFFF ? ph[k] : 0x7FFFFFFF);
    }
  }
}

// The walk kernel.  Without a queue, workgroup i walks item i (the
// dispatcher hands out items in list order).  With a queue (persistent
// form), a grid sized to the resident waves takes items from an atomic
// counter until the list is exhausted; every wave leaves when the counter
// passes n_walk.
// The 4- and 8-slot walks are held to 256 registers (two waves per SIMD):
// the 4-slot walk fits without spilling; the 8-slot walk spills, which its
// row bands more than make up for, since 4-slot waves can then share its
// SIMDs.  The 16-, 32- and 64-slot walks take what they need.
template <int K, int kMode>
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(K <= 8 ? 2 : 1)))
WalkKernel(WalkParams P) {
  extern __shared__ int32_t lds[];
  // The 8- and 16-slot walks hold the paths with the widest active lists,
  // the longest sequential walks of a frame: their waves take precedence in
  // instruction issue over the 4-slot waves sharing their SIMD.
  if (K >= 8) __builtin_amdgcn_s_setprio(2);
  // one call site, so the item body is inlined once
  auto take = [&]() -> int {
    if (!P.queue) return (int)blockIdx.x;
    int got = 0;
    if (Lane() == 0) got = atomicAdd(P.queue, 1);
    return __builtin_amdgcn_readlane(got, 0);
  };
  for (int wi = take(); wi < P.n_walk; wi = take()) {
    WalkItem<K, kMode>(P, wi, lds);
    if (!P.queue) break;
    WaveSync();
  }
}

// The 32- and 64-slot walks (active lists of up to 2048 and 4096 edges) are
// instantiated in their own translation unit, hipr_walk_wide.hip, which
// compiles beside the main one; they walk unbanded jobs in the count and
// write modes only.  Launches WalkKernel<K, mode> for K = 32 or 64.
hipError_t LaunchWideWalk(int K, int mode, unsigned grid, size_t lds, hipStream_t st,
                          const WalkParams& P);

#ifndef HIPR_TEMPLATES_ONLY
// Grid-stride fill of n int32 words.
__global__ void FillKernel(int32_t* __restrict__ p, int64_t n, int32_t v) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    p[i] = v;
}

// Zeroes the coverage records of the listed jobs (before a write pass redoes
// them).
__global__ void ResetCovKernel(const int32_t* __restrict__ list, int n, JobCov* __restrict__ cov) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) cov[list[i]] = JobCov{};
}
#endif  // HIPR_TEMPLATES_ONLY

}  // namespace hipr

#endif  // SKITY_HIP_WALK_HPP_
