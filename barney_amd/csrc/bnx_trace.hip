// bnx_trace.hip -- two-level closest-hit / any-hit traversal on gfx950.
//
// Replaces TraceInterface::traceRay (rtcore/cuda/TraceInterface.h:24-139,
// CPU: rtcore/cpu/Device.cpp:112-201) together with the trace kernel of
// kernels/traceRays.dev.cu:26-54.
//
// Design (wave64):
//   * queue traversal, one ray per thread (k_trace_static), or, for the
//     secondary rays of large scenes, from a per-segment pool (k_trace_pool):
//     a wave claims 64 rays at a time and refills lanes as their walks end,
//     so short walks do not idle behind the wave's longest;
//   * nearest-child-first traversal of a 4-wide BVH (collapsed from the
//     binary build) over 128-byte nodes: the four child boxes in one fetch
//     (seven 16-byte loads), entry distances ordered by a sorting network;
//   * short stack: a 16-entry ring per lane in LDS laid out [entry][lane]
//     (bank-conflict free), spilling its oldest entries to a per-lane
//     private array when deeper;
//   * two levels: a TLAS leaf pushes a LEAVE marker and switches the ray to
//     object space (enterBlas, rtcore/cuda/TraceInterface.h:107-122);
//   * primitive programs are a closed type switch (triangles / spheres)
//     instead of SBT function pointers (rtcore/cuda/Geom.h:21-36);
//   * ties in t go to the lower (instance, geometry, primitive) ID, and a
//     box whose entry distance equals the current best is still visited;
//   * the extended kernels (EXT) add the cylinder leaf test and the
//     renderer's cut plane, which rejects surface hits on its negative side
//     in every trace (OptixGlobals::hitOnInvisibleSide, render/
//     OptixGlobals.h:36-45); worlds without cylinders traced without a cut
//     plane keep the plain kernels and their register allocation.
#include "bnx_internal.h"
#include "bnx_render.h"

namespace bnx {
namespace {

constexpr int TB = 256;        // threads per block
// waves per SIMD the queue-trace kernels are compiled for (register budget:
// 512 / waves VGPRs); the walks are latency bound, so more resident waves
// hide more of each dependent node fetch
#ifndef BNX_TRACE_WAVES
#define BNX_TRACE_WAVES 7
#endif
// traversal stack: the top SH entries per lane in an LDS ring (SH a power
// of two: 8, 16 or 32, chosen per device; 16 by default), the rest spilled
// to a private array, STACK_CAP entries in all whatever SH is
constexpr int STACK_CAP = 192;
constexpr int API_SHORT = 16;  // the ray-array entry point's ring

struct HitRec {
  float t, u, v;
  int prim, geom, inst;
};

BNX_D bool id_less(int i0, int g0, int p0, int i1, int g1, int p1)
{
  if (i0 != i1) return i0 < i1;
  if (g0 != g1) return g0 < g1;
  return p0 < p1;
}

BNX_D bool accept_t(float t, const HitRec &b, int inst, int geom, int prim)
{
  if (t < b.t) return true;
  if (t > b.t) return false;
  return b.prim < 0 ? false : id_less(inst, geom, prim, b.inst, b.geom, b.prim);
}

// The cut plane (nx, ny, nz, w; off when w <= -1e28) and the ray's world
// frame (its queue record: origin at rorg, direction at rdir).  Every
// surface program of the reference rejects a hit at t whose world position
// org + t dir lies on the plane's negative side (OptixGlobals::
// hitOnInvisibleSide, render/OptixGlobals.h:36-45; Triangles.dev.cu:39,
// Spheres.dev.cu:170, Cylinders.dev.cu:258).  t is the same in the object
// frame of an instance, whose transform is affine.
//
// Stochastic opacity (EXT kernels): a triangle of a geometry whose alpha can
// be below 1 (GeomDev::translucent) is kept as a candidate only with
// probability alpha (Triangles.dev.cu:120-133).  Its draw is taken from a
// copy of the ray's RNG state advanced by the bits of the candidate's
// object-space position, so each candidate's fate depends on the ray and the
// candidate alone and the closest kept candidate is the same in any
// traversal order (the reference advances the ray's own state, in its
// traversal order).  rng null: every surface is opaque (the ray-array entry
// point).  The closest candidate the opacity test rejected goes to rej: the
// ID channels report the closest candidate whatever its opacity
// (Triangles.dev.cu:45-60).
struct Cut {
  float4 plane;
  const float4 *rorg, *rdir;
  const uint64_t *rng;
  const SceneDev *S;
  int geom_base;          // the walk's instance: its first geometry-table entry
  bool translucent;       // its BLAS holds a translucent geometry
  HitRec *rej;
};
BNX_D Cut no_cut()
{
  return Cut{make_float4(0.f, 0.f, 0.f, -1e30f), nullptr, nullptr, nullptr, nullptr, 0, false, nullptr};
}

// whether the candidate (triangle prim of geometry G at barycentrics u, v,
// rtc instance inst) survives its opacity draw
BNX_D bool opacity_keeps(const SceneDev &S, const GeomDev &G, int inst, int prim, float u, float v, uint64_t seed)
{
  const int *ix = G.indices + 3 * prim;
  const v3 v0 = ld3(G.vertices + 3 * ix[0]), v1 = ld3(G.vertices + 3 * ix[1]), v2 = ld3(G.vertices + 3 * ix[2]);
  const v3 osP = (((1.f - u) - v) * v0 + u * v1) + v * v2;
  const InstDev &I = S.insts[inst];
  const v3 P = I.identity ? osP : xfm_point_rows(I.o2w, osP);
  const float4 base = material_color(S, G, inst, prim, tri_corners(ix, prim, u, v), make_float4(0.f, 0.f, 0.f, 1.f),
                                     P, osP);
  const float alpha = material_alpha(G, base);
  if (alpha < 1.f) {
    rng_next_h(seed, (uint64_t)__float_as_uint(osP.x));
    rng_next_h(seed, (uint64_t)__float_as_uint(osP.y));
    rng_next_h(seed, (uint64_t)__float_as_uint(osP.z));
    Random2 rng(seed, 290374u);
    if (rng() > alpha) return false;
  }
  return true;
}
template <bool EXT> BNX_D bool cut_rejects(const Cut &c, float t)
{
  if (!EXT || c.plane.w <= -1e28f) return false;
  const v3 o = xyz(*c.rorg);
  const v3 d = fix_dir(xyz(*c.rdir));
  const v3 P = o + t * d;
  return dot(v3{c.plane.x, c.plane.y, c.plane.z}, P) + c.plane.w < 0.f;
}

// barney's triangle test (rtcore/hiprt/TraceInterface.h:38-76) on a BVH-
// ordered record (v0|prim, e1|geom, e2); acceptance under the tie rule.
// Evaluated without early exits: the reference's returns (a ray parallel to
// the plane, t behind tmin, a degenerate projection, a point outside the
// triangle) become one acceptance test after the three divisions, which
// keeps a wave's lanes together through the test and lets the three record
// loads issue at once (the values of every accepted hit are the same
// operations as the reference's, in the same order)
template <bool EXT>
BNX_D bool tri_test(const float4 *rec, v3 o, v3 d, float tmin, HitRec &best, int inst, const Cut &cut)
{
  const float4 A = rec[0], B = rec[1], C = rec[2];
  const v3 v0 = xyz(A), e1 = xyz(B), e2 = xyz(C);
  const int prim = __float_as_int(A.w), geom = __float_as_int(B.w);
  const v3 N = cross(e1, e2);
  const float dn = dot(d, N);
  const float t = -dot(o - v0, N) / dn;
  const v3 P = (o - v0) + t * d;
  const bool px = fabsf(N.x) >= fmx(fabsf(N.y), fabsf(N.z));
  const bool pz = !px && !(fabsf(N.y) > fabsf(N.z));
  const float e1u = px ? e1.y : e1.x, e2u = px ? e2.y : e2.x, Pu = px ? P.y : P.x;
  const float e1v = pz ? e1.y : e1.z, e2v = pz ? e2.y : e2.z, Pv = pz ? P.y : P.z;
  const float den = e1u * e2v - e1v * e2u;
  const float u = (Pu * e2v - Pv * e2u) / den;
  const float v = (e1u * Pv - e1v * Pu) / den;
  const bool in = !(fabsf(dn) < 1e-12f) & !(t <= tmin) & !(den == 0.f) & !(u < 0.f) & !(v < 0.f) & !((u + v) >= 1.f);
  if (!in || !accept_t(t, best, inst, geom, prim)) return false;
  if (cut_rejects<EXT>(cut, t)) return false;
  if (EXT && cut.translucent && cut.rng) {
    const GeomDev &G = cut.S->geoms[cut.geom_base + geom];
    if (G.translucent && !opacity_keeps(*cut.S, G, inst, prim, u, v, *cut.rng)) {
      HitRec &r = *cut.rej;
      if (t < r.t || (t == r.t && (r.prim < 0 || id_less(inst, geom, prim, r.inst, r.geom, r.prim))))
        r = HitRec{t, u, v, prim, geom, inst};
      return false;
    }
  }
  best = HitRec{t, u, v, prim, geom, inst};
  return true;
}

// SpheresPrograms::intersect (Spheres.dev.cu:114-166), "move the origin"
template <bool EXT>
BNX_D bool sphere_test(float4 S, const int2 *ref, v3 o, v3 d, float tmin, HitRec &best, int inst, const Cut &cut)
{
  v3 center = xyz(S);
  float radius = S.w;
  float t_move = fmx(0.f, length(center - o) - 3.f * radius);
  v3 om = o + t_move * d;
  float tmin_l = fmx(0.f, tmin - t_move);
  v3 oc = om - center;
  float a = dot(d, d);
  float b = dot(oc, d);
  float c = dot(oc, oc) - radius * radius;
  float disc = b * b - a * c;
  if (disc < 0.f) return false;
  float sq = sqrtf(disc);
  float r1 = (-b - sq) / a;
  float r2 = (-b + sq) / a;
  float cand;
  if (r1 > tmin_l) cand = r1;
  else if (r2 > tmin_l) cand = r2;
  else return false;
  float t = cand + t_move;
  if (t > best.t) return false;
  int2 r = *ref;
  if (!accept_t(t, best, inst, r.x, r.y)) return false;
  if (cut_rejects<EXT>(cut, t)) return false;
  best = HitRec{t, cand, 0.f, r.y, r.x, inst};
  return true;
}

// The user geometries with two records per primitive: cylinders
// (CylindersPrograms::intersect, native/geometry/Cylinders.dev.cu:107-260,
// records (v0, radius), (v1, 0)), capsules (CapsulesPrograms::intersect,
// Capsules.dev.cu:154-300, records (v0, r0), (v1, r1)) and cones
// (ConesPrograms::intersect, Cones.dev.cu:60-266, records (p0, ra),
// (p1, rb)).  The reference intersects them against the ray's closest hit
// so far; capsules and cones take only candidates strictly nearer, so they
// are given a bound just above it and a candidate at exactly the closest
// distance reaches the tie rule (the nearest candidate below the bound does
// not depend on the bound).  u keeps the hit distance for the shading.
BNX_D float tie_bound(float t) { return t * 1.0001f; }

// whether a capsule candidate at t survives its opacity draw (Capsules.dev.
// cu:266-280): the material's alpha from the hit attributes, the draw from
// Random(ray seed, hash(rtc instance, geometry, primitive)) taken on a copy
// of the ray's state, so that the candidate's fate depends on the ray and
// the candidate alone
BNX_D bool capsule_keeps(const SceneDev &S, const GeomDev &G, int inst, int geom, int prim, float4 A, float4 B, v3 o,
                         v3 d, float t, uint64_t seed)
{
  PairHit h;
  capsule_hit(S.insts[inst], A, B, o, d, t, h);
  const float4 base = material_color(S, G, inst, prim, pair_corners(G, prim, h.lerp), make_float4(0.f, 0.f, 0.f, 1.f),
                                     h.P, h.osP);
  const float alpha = material_alpha(G, base);
  if (alpha < 1.f) {
    const uint64_t hv = fnv3((uint32_t)inst, (uint32_t)geom, (uint32_t)prim);
    rng_next_h(seed, hv ? hv : 290374ull);
    const float r = (float)(rng_next(seed) & 0x00FFFFFFu) * (1.f / (float)0x01000000);
    if (r > alpha) return false;
  }
  return true;
}

BNX_D bool pair_test(int type, const float4 *rec, const int2 *ref, v3 o, v3 d, float tmin, HitRec &best, int inst,
                     const Cut &cut)
{
  const float4 A = rec[0], B = rec[1];
  float t;
  bool h;
  if (type == BLAS_CYLINDERS) h = cyl_intersect(o, d, xyz(A), xyz(B), A.w, tmin, best.t, t, nullptr);
  else if (type == BLAS_CAPSULES) h = capsule_intersect(o, d, A, B, tie_bound(best.t), t, nullptr, nullptr);
  else h = cone_intersect(o, d, A, B, tmin, tie_bound(best.t), t, nullptr, nullptr);
  if (!h) return false;
  const int2 r = *ref;
  if (!accept_t(t, best, inst, r.x, r.y)) return false;
  if (type == BLAS_CAPSULES && cut.translucent && cut.rng) {
    const GeomDev &G = cut.S->geoms[cut.geom_base + r.x];
    if (G.translucent && !capsule_keeps(*cut.S, G, inst, r.x, r.y, A, B, o, d, t, *cut.rng)) return false;
  }
  if (cut_rejects<true>(cut, t)) return false;
  best = HitRec{t, t, 0.f, r.y, r.x, inst};
  return true;
}

struct Counts {
  uint32_t nodes, tris, spheres;
};

// The ray in the world frame, read again from its record (origin at rorg,
// direction at rdir) when a walk leaves an instance, rather than held in
// registers for the whole walk: the walks are short of registers, and a
// ray changes frame only on entering and leaving an instance.
BNX_D void world_ray(const float4 *rorg, const float4 *rdir, v3 &o, v3 &d, v3 &inv)
{
  o = xyz(*rorg);
  d = fix_dir(xyz(*rdir));
  inv = v3{1.f / d.x, 1.f / d.y, 1.f / d.z};
}

// the ray in an instance's object frame (identity instances keep the world
// frame the walk is in)
BNX_D void instance_ray(const InstDev &I, const float4 *rorg, const float4 *rdir, v3 &o, v3 &d, v3 &inv)
{
  if (I.identity) return;
  o = xfm_point_rows(I.w2o, xyz(*rorg));
  d = fix_dir(xfm_vector_rows(I.w2o, fix_dir(xyz(*rdir))));
  inv = v3{1.f / d.x, 1.f / d.y, 1.f / d.z};
}

// entry distance of child k of a 4-wide node, +inf when the ray misses it.
// Branch-free, so the six box vectors of a node are loaded together up front
// (a short-circuit on the child code let the compiler sink each child's box
// loads behind the code load, doubling the dependent fetches per node).
// |inv| <= 1e6 (|d| >= 1e-6 after fix_dir) and empty slots hold inverted
// infinite boxes, so no slab distance is NaN and fminf/fmaxf give the same
// values as compares.
BNX_D float child_entry(v3 o, v3 inv, float lx, float hx, float ly, float hy, float lz, float hz, uint32_t code,
                        float tmin, float tmax)
{
  const float t0x = (lx - o.x) * inv.x, t1x = (hx - o.x) * inv.x;
  const float t0y = (ly - o.y) * inv.y, t1y = (hy - o.y) * inv.y;
  const float t0z = (lz - o.z) * inv.z, t1z = (hz - o.z) * inv.z;
  const float tn = fmaxf(fmaxf(fmaxf(tmin, fminf(t0x, t1x)), fminf(t0y, t1y)), fminf(t0z, t1z));
  const float tf = fminf(fminf(fminf(tmax, fmaxf(t0x, t1x)), fmaxf(t0y, t1y)), fmaxf(t0z, t1z));
  const bool h = (tn <= tf * 1.00001f) & (code != EMPTY4);
  return h ? tn : __builtin_huge_valf();
}


BNX_D void cswap(float &ta, uint32_t &ca, float &tb, uint32_t &cb)
{
  const bool s = tb < ta;
  const float t = s ? tb : ta;
  const uint32_t c = s ? cb : ca;
  tb = s ? ta : tb;
  cb = s ? ca : cb;
  ta = t;
  ca = c;
}

// lds: this lane's ring, &ring[0][lane]; the stack depth and the spill
// array are locals so the depth stays in a register.  XF: some instance
// transforms the ray (a non-identity instance in a launched world); without
// it the frame never changes, so the ray (o, d, inv) is loop-invariant and
// the compiler carries no copies of it around the walk's paths.
template <int SH, bool EXT, bool XF>
BNX_D HitRec traverse(const SceneDev &S, const float4 *rorg, const float4 *rdir, float tmin, float tmax, bool any_hit,
                      uint32_t *lds, Counts &cnt, float4 cut_plane, const uint64_t *rng, HitRec &rej)
{
  Cut cut{cut_plane, rorg, rdir, rng, &S, 0, false, &rej};
  uint32_t deep[STACK_CAP - SH];
  int sp = 0;
  auto push = [&](uint32_t v) {
    const int s = sp & (SH - 1);
    if (sp >= SH) deep[sp - SH] = lds[s * TB];
    lds[s * TB] = v;
    sp++;
  };
  auto pop = [&]() -> uint32_t {
    sp--;
    const int s = sp & (SH - 1);
    const uint32_t v = lds[s * TB];
    if (sp >= SH) lds[s * TB] = deep[sp - SH];
    return v;
  };
  HitRec best{tmax, 0.f, 0.f, -1, -1, -1};
  v3 o, d, inv;
  world_ray(rorg, rdir, o, d, inv);
  const Node4 *nodes = S.tlas_nodes;
  BlasDev blas{};
  bool in_blas = false;
  int inst = -1;
  uint32_t cur = 0;
  uint32_t nvis = 0;
  // every node is entered at most once per ray, so a ray can take no more
  // loop trips than the scene has nodes and leaves; the cap only bounds the
  // loop if the structure were corrupted
  for (uint32_t trip = 0; trip < (1u << 28); trip++) {
    if (cur < EMPTY4) {
      // ------------------------------------------------ 4-wide inner node
      nvis++;
      const float4 *np = (const float4 *)(nodes + cur);
      const float4 lx = np[0], hx = np[1], ly = np[2], hy = np[3], lz = np[4], hz = np[5];
      const uint4 ch = ((const uint4 *)np)[6];
      float t0 = child_entry(o, inv, lx.x, hx.x, ly.x, hy.x, lz.x, hz.x, ch.x, tmin, best.t);
      float t1 = child_entry(o, inv, lx.y, hx.y, ly.y, hy.y, lz.y, hz.y, ch.y, tmin, best.t);
      float t2 = child_entry(o, inv, lx.z, hx.z, ly.z, hy.z, lz.z, hz.z, ch.z, tmin, best.t);
      float t3 = child_entry(o, inv, lx.w, hx.w, ly.w, hy.w, lz.w, hz.w, ch.w, tmin, best.t);
      uint32_t c0 = ch.x, c1 = ch.y, c2 = ch.z, c3 = ch.w;
      // nearest first: a 4-element sorting network on the entry distances
      cswap(t0, c0, t1, c1);
      cswap(t2, c2, t3, c3);
      cswap(t0, c0, t2, c2);
      cswap(t1, c1, t3, c3);
      cswap(t1, c1, t2, c2);
      const float INF = __builtin_huge_valf();
      if (t0 < INF) {
        if (t3 < INF) push(c3);
        if (t2 < INF) push(c2);
        if (t1 < INF) push(c1);
        cur = c0;
        continue;
      }
    } else if (cur == LEAVE_BLAS) {
      in_blas = false;
      if (XF) world_ray(rorg, rdir, o, d, inv);
      nodes = S.tlas_nodes;
    } else {
      // ------------------------------------------------------------ leaf
      uint32_t first = (cur & ~LEAF_BIT) >> 3;
      int count = (int)(cur & 7u);
      if (!in_blas) {
        inst = S.tlas_order[first];                  // TLAS leaves hold one instance
        const InstDev &I = S.insts[inst];
        blas = S.blas[I.blas];
        if (EXT) {
          cut.geom_base = I.geom_base;
          cut.translucent = blas.translucent != 0;
        }
        if (XF && !I.identity) instance_ray(I, rorg, rdir, o, d, inv);
        push(LEAVE_BLAS);
        in_blas = true;
        nodes = blas.nodes;
        cur = 0;
        continue;
      }
      bool found = false;
      if (blas.type == BLAS_TRIANGLES) {
        cnt.tris += count;
        for (int k = 0; k < count; k++)
          found |= tri_test<EXT>(blas.prims + 3 * (first + k), o, d, tmin, best, inst, cut);
      } else if (!EXT || blas.type == BLAS_SPHERES) {
        cnt.spheres += count;
        for (int k = 0; k < count; k++)
          found |= sphere_test<EXT>(blas.prims[first + k], blas.refs + first + k, o, d, tmin, best, inst, cut);
      } else {
        cnt.spheres += count;
        for (int k = 0; k < count; k++)
          found |= pair_test(blas.type, blas.prims + 2 * (first + k), blas.refs + first + k, o, d, tmin, best, inst,
                               cut);
      }
      if (any_hit && found) break;
    }
    if (sp == 0) break;
    cur = pop();
  }
  cnt.nodes += nvis;
  return best;
}

// The walk of traverse() in speculative while-while order (Aila and Laine,
// "Understanding the Efficiency of Ray Traversal on GPUs", HPG 2009): a
// lane that reaches a primitive leaf sets it aside and keeps stepping inner
// nodes from its stack while other lanes of the wave still step nodes, until
// every lane with a node to step already holds a leaf (or none has a node);
// then the lanes test their leaves together.  Each trip of the wave thus
// serves either node steps or leaf tests for most of its lanes, where the
// plain walk pays for both whenever its lanes are split between them.  A
// node is culled against the best hit known when it is stepped, so a lane
// may step nodes that a hit in its pending leaf would have culled; the hit
// returned is the same (equal t resolves by ID).  cur == EMPTY4: the stack
// is exhausted.
template <int SH, bool EXT, bool XF>
BNX_D HitRec traverse_spec(const SceneDev &S, const float4 *rorg, const float4 *rdir, float tmin, float tmax,
                           bool any_hit, uint32_t *lds, Counts &cnt, float4 cut_plane, const uint64_t *rng,
                           HitRec &rej)
{
  Cut cut{cut_plane, rorg, rdir, rng, &S, 0, false, &rej};
  uint32_t deep[STACK_CAP - SH];
  int sp = 0;
  auto push = [&](uint32_t v) {
    const int s = sp & (SH - 1);
    if (sp >= SH) deep[sp - SH] = lds[s * TB];
    lds[s * TB] = v;
    sp++;
  };
  auto pop = [&]() -> uint32_t {
    if (sp == 0) return EMPTY4;
    sp--;
    const int s = sp & (SH - 1);
    const uint32_t v = lds[s * TB];
    if (sp >= SH) lds[s * TB] = deep[sp - SH];
    return v;
  };
  HitRec best{tmax, 0.f, 0.f, -1, -1, -1};
  v3 o, d, inv;
  world_ray(rorg, rdir, o, d, inv);
  const Node4 *nodes = S.tlas_nodes;
  BlasDev blas{};
  bool in_blas = false;
  int inst = -1;
  uint32_t cur = 0;
  uint32_t leaf = 0;             // primitive leaf set aside (leaf codes carry LEAF_BIT)
  uint32_t nvis = 0;
  // every node is entered at most once per ray; the cap only bounds the
  // loops if the structure were corrupted
  uint32_t trips = 0;
  while (trips < (1u << 28)) {
    // ------------------------------------------------------- node phase
    for (;; trips++) {
      if (leaf == 0 && in_blas && (cur & LEAF_BIT)) {
        leaf = cur;
        cur = pop();
      }
      const bool step = cur < EMPTY4;
      if (!__any(step && leaf == 0) || trips >= (1u << 28)) break;
      if (step) {
        const float4 *np = (const float4 *)(nodes + cur);
        const float4 lx = np[0], hx = np[1], ly = np[2], hy = np[3], lz = np[4], hz = np[5];
        const uint4 ch = ((const uint4 *)np)[6];
        nvis++;
        float t0 = child_entry(o, inv, lx.x, hx.x, ly.x, hy.x, lz.x, hz.x, ch.x, tmin, best.t);
        float t1 = child_entry(o, inv, lx.y, hx.y, ly.y, hy.y, lz.y, hz.y, ch.y, tmin, best.t);
        float t2 = child_entry(o, inv, lx.z, hx.z, ly.z, hy.z, lz.z, hz.z, ch.z, tmin, best.t);
        float t3 = child_entry(o, inv, lx.w, hx.w, ly.w, hy.w, lz.w, hz.w, ch.w, tmin, best.t);
        uint32_t c0 = ch.x, c1 = ch.y, c2 = ch.z, c3 = ch.w;
        cswap(t0, c0, t1, c1);
        cswap(t2, c2, t3, c3);
        cswap(t0, c0, t2, c2);
        cswap(t1, c1, t3, c3);
        cswap(t1, c1, t2, c2);
        const float INF = __builtin_huge_valf();
        if (t0 < INF) {
          if (t3 < INF) push(c3);
          if (t2 < INF) push(c2);
          if (t1 < INF) push(c1);
          cur = c0;
        } else {
          cur = pop();
        }
      }
    }
    // ------------------------------------------------------- leaf phase
    bool found = false;
    if (leaf != 0) {
      const uint32_t first = (leaf & ~LEAF_BIT) >> 3;
      const int count = (int)(leaf & 7u);
      if (blas.type == BLAS_TRIANGLES) {
        cnt.tris += count;
        for (int k = 0; k < count; k++)
          found |= tri_test<EXT>(blas.prims + 3 * (first + k), o, d, tmin, best, inst, cut);
      } else if (!EXT || blas.type == BLAS_SPHERES) {
        cnt.spheres += count;
        for (int k = 0; k < count; k++)
          found |= sphere_test<EXT>(blas.prims[first + k], blas.refs + first + k, o, d, tmin, best, inst, cut);
      } else {
        cnt.spheres += count;
        for (int k = 0; k < count; k++)
          found |= pair_test(blas.type, blas.prims + 2 * (first + k), blas.refs + first + k, o, d, tmin, best, inst,
                               cut);
      }
      leaf = 0;
    } else if ((cur & LEAF_BIT) && !in_blas) {
      // a TLAS leaf (the node phase sets every BLAS leaf aside): enter its
      // instance
      inst = S.tlas_order[(cur & ~LEAF_BIT) >> 3];
      const InstDev &I = S.insts[inst];
      blas = S.blas[I.blas];
      if (EXT) {
        cut.geom_base = I.geom_base;
        cut.translucent = blas.translucent != 0;
      }
      if (XF && !I.identity) {
        instance_ray(I, rorg, rdir, o, d, inv);
      }
      push(LEAVE_BLAS);
      in_blas = true;
      nodes = blas.nodes;
      cur = 0;
    } else if (cur == LEAVE_BLAS) {
      in_blas = false;
      if (XF) world_ray(rorg, rdir, o, d, inv);
      nodes = S.tlas_nodes;
      cur = pop();
    }
    if ((any_hit && found) || (cur == EMPTY4 && leaf == 0)) break;
    trips++;
  }
  cnt.nodes += nvis;
  return best;
}

__global__ __launch_bounds__(TB) void k_trace_api(SceneDev S, const bnx_ray *rays, int n, bnx_hit *hits,
                                                  uint32_t flags)
{
  __shared__ uint32_t ring[API_SHORT * TB];
  uint32_t *lds = ring + threadIdx.x;
  {
    const int i = blockIdx.x * TB + threadIdx.x;
    if (i >= n) return;
    const float4 *rp = (const float4 *)(rays + i);
    float4 r0 = rp[0], r1 = rp[1];
    Counts cnt{0, 0, 0};
    HitRec rej{__builtin_huge_valf(), 0.f, 0.f, -1, -1, -1};
    HitRec h = traverse<API_SHORT, true, true>(S, rp, rp + 1, r0.w, r1.w, (flags & BNX_TRACE_ANY) != 0, lds, cnt,
                                               no_cut().plane, nullptr, rej);
    bnx_hit out;
    out.t = h.t;
    out.u = h.prim >= 0 ? h.u : 0.f;
    out.v = h.prim >= 0 ? h.v : 0.f;
    out.prim_id = h.prim;
    if (h.prim >= 0) {
      const InstDev &I = S.insts[h.inst];
      out.geom_id = h.geom;
      out.inst = h.inst;
      out.user_inst_id = I.user_id;
      out.obj_id = S.geoms[I.geom_base + h.geom].user_id;
    } else {
      out.geom_id = out.inst = out.user_inst_id = out.obj_id = -1;
    }
    hits[i] = out;
  }
}

// A translucent candidate the opacity test rejected lies in front of the
// ray's hit: for a primary ray of a request's first sample the trace writes
// the ID channels from it (as the reference's any-hit program writes them
// before the opacity test, Triangles.dev.cu:45-60) and flags the ray so that
// shade leaves them.
BNX_D void translucent_ids(const SlotDev &P, const TraceQueue &q, int i, float4 d, const HitRec &rej)
{
  const uint4 ps = q.ps[i];
  const uint32_t f = __float_as_uint(d.w);
  if ((ps.y >> 24) != 0 || (f & FLAG_SHADOW) || P.accum_base + flag_sample(f) != 0) return;
  const int pixel = (int)ps.z;
  const InstDev &I = P.scene.insts[rej.inst];
  P.fb.prim[pixel] = rej.prim;
  P.fb.inst[pixel] = I.user_id;
  P.fb.obj[pixel] = P.scene.geoms[I.geom_base + rej.geom].user_id;
  q.dir[i].w = __uint_as_float(f | FLAG_IDS);
}

// one ray per thread over the generation's segments (256-thread block map);
// each block lies inside one request's segment, so the request's scene
// pointers are block-uniform.  counters: statistics array (STAT_NODES,
// STAT_TRIS, STAT_SPHERES)
template <int SH, bool SPEC, bool EXT, bool XF>
BNX_D void trace_static(const SlotDev *__restrict__ slots, const SegView &sv, const TraceQueue &q,
                        unsigned long long *counters)
{
  __shared__ uint32_t ring[SH * TB];
  uint32_t *lds = ring + threadIdx.x;
  Counts cnt{0, 0, 0};
  int i;
  bool valid;
  const Segment sg = segment_of<TB>(sv, i, valid);
  const SceneDev S = slots[sg.slot].scene;
  const float4 cut_plane = EXT ? slot_cut_plane(slots[sg.slot]) : no_cut().plane;
  if (valid) {
    float4 o = q.org[i];
    float4 d = q.dir[i];
    const bool shadow = (__float_as_uint(d.w) & FLAG_SHADOW) != 0;
    HitRec h{o.w, 0.f, 0.f, -1, -1, -1};
    HitRec rej{__builtin_huge_valf(), 0.f, 0.f, -1, -1, -1};
    if (S.num_insts > 0)
      h = SPEC ? traverse_spec<SH, EXT, XF>(S, q.org + i, q.dir + i, 0.f, o.w, shadow, lds, cnt, cut_plane,
                                            q.rng + i, rej)
               : traverse<SH, EXT, XF>(S, q.org + i, q.dir + i, 0.f, o.w, shadow, lds, cnt, cut_plane, q.rng + i,
                                       rej);
    if (h.prim >= 0) q.org[i].w = h.t;
    q.hit[i] = make_int4(h.prim, h.geom, h.inst, __float_as_int(h.u));
    q.hit_v[i] = h.v;
    if (EXT && rej.prim >= 0 && rej.t < h.t) translucent_ids(slots[sg.slot], q, i, d, rej);
  }
  uint32_t a = wave_sum(cnt.nodes), b = wave_sum(cnt.tris), c = wave_sum(cnt.spheres);
  if (lane_id() == 0) {
    stat_add(counters, STAT_NODES, a);
    stat_add(counters, STAT_TRIS, b);
    stat_add(counters, STAT_SPHERES, c);
  }
}

// EXT: the leaf test knows cylinder BLASes and the cut plane of each
// request.  The kernels without it serve worlds of triangles and spheres
// traced without a cut plane; the launcher takes the _ext kernels while any
// built world holds cylinders or a request in the launch has a cut plane.
template <int SH, bool SPEC, bool XF>
__global__ __launch_bounds__(TB) __attribute__((amdgpu_waves_per_eu(BNX_TRACE_WAVES))) void k_trace_static(
    const SlotDev *__restrict__ slots, SegView sv, TraceQueue q, unsigned long long *counters)
{
  trace_static<SH, SPEC, false, XF>(slots, sv, q, counters);
}

template <int SH, bool SPEC, bool XF>
__global__ __launch_bounds__(TB) __attribute__((amdgpu_waves_per_eu(BNX_TRACE_WAVES))) void k_trace_static_ext(
    const SlotDev *__restrict__ slots, SegView sv, TraceQueue q, unsigned long long *counters)
{
  trace_static<SH, SPEC, true, XF>(slots, sv, q, counters);
}

// The generation's rays, segment by segment: every block lies inside one
// request's segment (its scene pointers are block-uniform) and draws rays
// from the segment's pool.  A wave claims POOL_CHUNK ray indices at a time
// (one atomic per chunk) and hands them to its lanes as they finish, so a
// lane whose walk ends early takes the next ray instead of idling until the
// wave's longest walk is over.  The walk is the one of traverse() above,
// taken one node or leaf per loop trip.
constexpr int POOL_CHUNK = 64;

template <int SH, bool EXT, bool XF>
BNX_D void trace_pool(const SlotDev *__restrict__ slots, const SegView &sv, const TraceQueue &q,
                      unsigned long long *counters)
{
  __shared__ uint32_t ring[SH * TB];
  uint32_t *lds = ring + threadIdx.x;
  uint32_t deep[STACK_CAP - SH];
  int sp = 0;
  auto push = [&](uint32_t v) {
    const int s = sp & (SH - 1);
    if (sp >= SH) deep[sp - SH] = lds[s * TB];
    lds[s * TB] = v;
    sp++;
  };
  auto pop = [&]() -> uint32_t {
    sp--;
    const int s = sp & (SH - 1);
    const uint32_t v = lds[s * TB];
    if (sp >= SH) lds[s * TB] = deep[sp - SH];
    return v;
  };
  const int seg = sv.blkT[blockIdx.x];
  const Segment sg = sv.segs[seg];
  const SceneDev S = slots[sg.slot].scene;
  const float4 cut_plane = EXT ? slot_cut_plane(slots[sg.slot]) : no_cut().plane;
  const bool empty_scene = S.num_insts == 0;
  const int refill_at = sg.refill;
  Counts cnt{0, 0, 0};
  // per-lane walk state
  bool active = false, any_hit = false, in_blas = false;
  int ray = 0, inst = -1;
  v3 o{0.f, 0.f, 0.f}, d{1.f, 1.f, 1.f}, inv{1.f, 1.f, 1.f};
  HitRec best{0.f, 0.f, 0.f, -1, -1, -1};
  HitRec rej{0.f, 0.f, 0.f, -1, -1, -1};
  int geom_base = 0;     // EXT: the lane's instance's first geometry-table entry
  const Node4 *nodes = S.tlas_nodes;
  BlasDev blas{};
  uint32_t cur = 0;
  // wave-uniform chunk [q_next, q_end) of segment-local ray indices: first
  // the wave's own chunk of the segment's static part, then chunks from the
  // pool (whose counter starts where the static part ends)
  int q_next = (((int)blockIdx.x - sg.blockT) * (TB / 64) + (int)(threadIdx.x >> 6)) * POOL_CHUNK;
  int q_end = min(q_next + POOL_CHUNK, sg.count);
  bool dry = false;
  for (;;) {
    // ------------------------------------------------------- refill
    const bool need = !active;
    const uint64_t mask = __ballot(need);
    if (mask != 0 && !dry) {
      if (q_next >= q_end) {
        int base = 0;
        if (lane_id() == 0) base = atomicAdd(sv.pool + seg, POOL_CHUNK);
        base = __builtin_amdgcn_readfirstlane(base);
        if (base >= sg.count) {
          dry = true;
          q_next = q_end = 0;
        } else {
          q_next = base;
          q_end = min(base + POOL_CHUNK, sg.count);
        }
      }
      const int avail = max(0, q_end - q_next);
      const int rank = __popcll(mask & lanemask_lt());
      if (need && rank < avail) {
        ray = sg.begin + q_next + rank;
        const float4 ro = q.org[ray], rd = q.dir[ray];
        any_hit = (__float_as_uint(rd.w) & FLAG_SHADOW) != 0;
        o = xyz(ro);
        d = fix_dir(xyz(rd));
        inv = v3{1.f / d.x, 1.f / d.y, 1.f / d.z};
        best = HitRec{ro.w, 0.f, 0.f, -1, -1, -1};
        rej = HitRec{__builtin_huge_valf(), 0.f, 0.f, -1, -1, -1};
        nodes = S.tlas_nodes;
        in_blas = false;
        inst = -1;
        cur = 0;
        sp = 0;
        active = true;
      }
      q_next += min(__popcll(mask), avail);
    }
    if (__ballot(active) == 0) break;
    // ---------------------------------------- walk until enough lanes idle
    while (active) {
      bool done = false;
      if (empty_scene) {
        done = true;
      } else if (cur < EMPTY4) {
        const float4 *np = (const float4 *)(nodes + cur);
        const float4 lx = np[0], hx = np[1], ly = np[2], hy = np[3], lz = np[4], hz = np[5];
        const uint4 ch = ((const uint4 *)np)[6];
        cnt.nodes++;
        float t0 = child_entry(o, inv, lx.x, hx.x, ly.x, hy.x, lz.x, hz.x, ch.x, 0.f, best.t);
        float t1 = child_entry(o, inv, lx.y, hx.y, ly.y, hy.y, lz.y, hz.y, ch.y, 0.f, best.t);
        float t2 = child_entry(o, inv, lx.z, hx.z, ly.z, hy.z, lz.z, hz.z, ch.z, 0.f, best.t);
        float t3 = child_entry(o, inv, lx.w, hx.w, ly.w, hy.w, lz.w, hz.w, ch.w, 0.f, best.t);
        uint32_t c0 = ch.x, c1 = ch.y, c2 = ch.z, c3 = ch.w;
        cswap(t0, c0, t1, c1);
        cswap(t2, c2, t3, c3);
        cswap(t0, c0, t2, c2);
        cswap(t1, c1, t3, c3);
        cswap(t1, c1, t2, c2);
        const float INF = __builtin_huge_valf();
        if (t0 < INF) {
          if (t3 < INF) push(c3);
          if (t2 < INF) push(c2);
          if (t1 < INF) push(c1);
          cur = c0;
        } else if (sp == 0) {
          done = true;
        } else {
          cur = pop();
        }
      } else if (cur == LEAVE_BLAS) {
        in_blas = false;
        if (XF) world_ray(q.org + ray, q.dir + ray, o, d, inv);
        nodes = S.tlas_nodes;
        if (sp == 0) done = true;
        else cur = pop();
      } else {
        const uint32_t first = (cur & ~LEAF_BIT) >> 3;
        const int count = (int)(cur & 7u);
        if (!in_blas) {
          inst = S.tlas_order[first];                  // TLAS leaves hold one instance
          const InstDev &I = S.insts[inst];
          blas = S.blas[I.blas];
          geom_base = I.geom_base;
          if (XF && !I.identity) {
            instance_ray(I, q.org + ray, q.dir + ray, o, d, inv);
          }
          push(LEAVE_BLAS);
          in_blas = true;
          nodes = blas.nodes;
          cur = 0;
        } else {
          bool found = false;
          const Cut cut{cut_plane, q.org + ray, q.dir + ray, q.rng + ray, &S, geom_base, blas.translucent != 0, &rej};
          if (blas.type == BLAS_TRIANGLES) {
            cnt.tris += count;
            for (int k = 0; k < count; k++)
              found |= tri_test<EXT>(blas.prims + 3 * (first + k), o, d, 0.f, best, inst, cut);
          } else if (!EXT || blas.type == BLAS_SPHERES) {
            cnt.spheres += count;
            for (int k = 0; k < count; k++)
              found |= sphere_test<EXT>(blas.prims[first + k], blas.refs + first + k, o, d, 0.f, best, inst, cut);
          } else {
            cnt.spheres += count;
            for (int k = 0; k < count; k++)
              found |= pair_test(blas.type, blas.prims + 2 * (first + k), blas.refs + first + k, o, d, 0.f, best,
                                   inst, cut);
          }
          if ((any_hit && found) || sp == 0) done = true;
          else cur = pop();
        }
      }
      if (done) {
        if (best.prim >= 0) q.org[ray].w = best.t;
        q.hit[ray] = make_int4(best.prim, best.geom, best.inst, __float_as_int(best.u));
        q.hit_v[ray] = best.v;
        if (EXT && rej.prim >= 0 && rej.t < best.t) translucent_ids(slots[sg.slot], q, ray, q.dir[ray], rej);
        active = false;
      }
      // leave the walk for a refill once refill_at lanes are idle (the test
      // is wave-uniform, so the lanes still walking leave together)
      if (!dry && 64 - __popcll(__ballot(true)) >= refill_at) break;
    }
  }
  uint32_t a = wave_sum(cnt.nodes), b = wave_sum(cnt.tris), c = wave_sum(cnt.spheres);
  if (lane_id() == 0) {
    stat_add(counters, STAT_NODES, a);
    stat_add(counters, STAT_TRIS, b);
    stat_add(counters, STAT_SPHERES, c);
  }
}

template <int SH, bool XF>
__global__ __launch_bounds__(TB) __attribute__((amdgpu_waves_per_eu(BNX_TRACE_WAVES))) void k_trace_pool(
    const SlotDev *__restrict__ slots, SegView sv, TraceQueue q, unsigned long long *counters)
{
  trace_pool<SH, false, XF>(slots, sv, q, counters);
}

template <int SH, bool XF>
__global__ __launch_bounds__(TB) __attribute__((amdgpu_waves_per_eu(BNX_TRACE_WAVES))) void k_trace_pool_ext(
    const SlotDev *__restrict__ slots, SegView sv, TraceQueue q, unsigned long long *counters)
{
  trace_pool<SH, true, XF>(slots, sv, q, counters);
}

template <int SH, bool SPEC, bool XF>
void launch_static_xf(Device *dev, bool ext, int blocks, const SlotDev *slots, const SegView &sv,
                      const TraceQueue &q, unsigned long long *counters)
{
  if (ext) k_trace_static_ext<SH, SPEC, XF><<<blocks, TB, 0, dev->stream>>>(slots, sv, q, counters);
  else k_trace_static<SH, SPEC, XF><<<blocks, TB, 0, dev->stream>>>(slots, sv, q, counters);
}

template <int SH, bool SPEC>
void launch_static(Device *dev, bool ext, bool xf, int blocks, const SlotDev *slots, const SegView &sv,
                   const TraceQueue &q, unsigned long long *counters)
{
  if (xf) launch_static_xf<SH, SPEC, true>(dev, ext, blocks, slots, sv, q, counters);
  else launch_static_xf<SH, SPEC, false>(dev, ext, blocks, slots, sv, q, counters);
}

template <int SH>
void launch_pool(Device *dev, bool ext, bool xf, int blocks, const SlotDev *slots, const SegView &sv,
                 const TraceQueue &q, unsigned long long *counters)
{
  if (ext) {
    if (xf) k_trace_pool_ext<SH, true><<<blocks, TB, 0, dev->stream>>>(slots, sv, q, counters);
    else k_trace_pool_ext<SH, false><<<blocks, TB, 0, dev->stream>>>(slots, sv, q, counters);
  } else {
    if (xf) k_trace_pool<SH, true><<<blocks, TB, 0, dev->stream>>>(slots, sv, q, counters);
    else k_trace_pool<SH, false><<<blocks, TB, 0, dev->stream>>>(slots, sv, q, counters);
  }
}

int grid_for(int n) { return n > 0 ? (n + TB - 1) / TB : 1; }

} // namespace

int trace_stack_capacity() { return STACK_CAP; }

int launch_trace_api(Group *g, const bnx_ray *d_rays, int n, bnx_hit *d_hits, uint32_t flags)
{
  Device *dev = g->dev;
  k_trace_api<<<grid_for(n), TB, 0, dev->stream>>>(g->scene, d_rays, n, d_hits, flags);
  BNX_HIP(hipGetLastError());
  return BNX_OK;
}

int launch_trace_queue(Device *dev, const SlotDev *slots, const SegView &sv, int blocks256, int blocksT, bool pooled,
                       bool cut, const TraceQueue &q, unsigned long long *counters)
{
  const int sh = dev->trace_stack_lds;
  const bool ext = ext_kernels(dev, cut);
  const bool xf = dev->xf_worlds > 0;
  if (pooled) {
    if (blocksT > 0) {
      if (sh == 8) launch_pool<8>(dev, ext, xf, blocksT, slots, sv, q, counters);
      else if (sh == 32) launch_pool<32>(dev, ext, xf, blocksT, slots, sv, q, counters);
      else launch_pool<16>(dev, ext, xf, blocksT, slots, sv, q, counters);
    }
  } else {
    if (blocks256 > 0) {
      const bool spec = dev->trace_walk == 1;
      if (sh == 8) {
        if (spec) launch_static<8, true>(dev, ext, xf, blocks256, slots, sv, q, counters);
        else launch_static<8, false>(dev, ext, xf, blocks256, slots, sv, q, counters);
      } else if (sh == 32) {
        if (spec) launch_static<32, true>(dev, ext, xf, blocks256, slots, sv, q, counters);
        else launch_static<32, false>(dev, ext, xf, blocks256, slots, sv, q, counters);
      } else {
        if (spec) launch_static<16, true>(dev, ext, xf, blocks256, slots, sv, q, counters);
        else launch_static<16, false>(dev, ext, xf, blocks256, slots, sv, q, counters);
      }
    }
  }
  BNX_HIP(hipGetLastError());
  return BNX_OK;
}

} // namespace bnx
