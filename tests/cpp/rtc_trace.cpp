// rtc_trace.cpp -- a scene built and traced through the C++ mirror of
// barney's rtc interface (include/bnx_rtc.hpp), the way barney's native
// layer drives a backend: createBuffer/upload, GeomType::createGeom,
// setVertices/setIndices, createTrianglesGroup/createUserGeomsGroup,
// createInstanceGroup, buildAccel, a trace launch, then setTransforms +
// refitAccel and a second trace.  Spheres and cylinders are user
// geometries (setSpheres / setCylinders, the Spheres::DD and Cylinders::DD
// fields).
//
// A third phase re-commits the triangles as Triangles::commit does
// (native/geometry/Triangles.cpp:50-80): the vertex buffer is uploaded
// again with every vertex moved, the same rtc::Geoms get setVertices /
// setIndices again, a new triangles group and a new instance group (all
// identity) are built, and the rays are traced again while the first
// groups still exist.  The rays and hits live in memory from the device's
// allocMem / allocHost and move with copyAsync / memsetAsync.
//
// usage: rtc_trace <in> <out>
//   in:  int32 nv, float3[nv], int32 nt, int3[nt], float3[nt] colours,
//        int32 ns, float3[ns] centres, float[ns] radii, float3 move,
//        int32 ncv, float3[ncv] cylinder vertices, int32 nc, int2[nc]
//        end points, float[nc] radii, int32 nr, bnx_ray[nr]
//   out: bnx_hit[nr] before the move, bnx_hit[nr] after it, bnx_hit[nr]
//        after the re-commit of moved vertices
// The triangles form two geometries (the first and second half), the
// spheres one, the cylinders one; instance 0 holds the spheres' group,
// instance 1 the triangles' group, instance 2 the cylinders' group, all
// with user instance ID 0; the move translates instance 1.  Prints "rtc_trace ok" and exits 0 on success.
#include "bnx_rtc.hpp"

#include <cstdio>
#include <fstream>
#include <iostream>
#include <vector>

using namespace bnx::rtc;

template <typename T> static std::vector<T> read_vec(std::ifstream &f, size_t n)
{
  std::vector<T> v(n);
  f.read((char *)v.data(), sizeof(T) * n);
  if (!f) throw std::runtime_error("rtc_trace: short input");
  return v;
}

static int32_t read_i32(std::ifstream &f) { return read_vec<int32_t>(f, 1)[0]; }

int main(int argc, char **argv)
{
  if (argc != 3) {
    std::cerr << "usage: rtc_trace <in> <out>" << std::endl;
    return 2;
  }
  try {
    std::ifstream in(argv[1], std::ios::binary);
    const int nv = read_i32(in);
    auto verts = read_vec<vec3f>(in, nv);
    const int nt = read_i32(in);
    auto idx = read_vec<vec3i>(in, nt);
    auto cols = read_vec<vec3f>(in, nt);
    const int ns = read_i32(in);
    auto centers = read_vec<vec3f>(in, ns);
    auto radii = read_vec<float>(in, ns);
    auto move = read_vec<vec3f>(in, 1)[0];
    const int ncv = read_i32(in);
    auto cverts = read_vec<vec3f>(in, ncv);
    const int nc = read_i32(in);
    auto cidx = read_vec<int32_t>(in, 2 * (size_t)nc);
    auto cradii = read_vec<float>(in, nc);
    const int nr = read_i32(in);
    auto rays = read_vec<bnx_ray>(in, nr);

    if (physicalDeviceCount() < 1) throw std::runtime_error("rtc_trace: no GPU");
    Device dev(0);
    GeomType *triType = dev.createGeomType(GeomType::TRIANGLES);
    GeomType *sphType = dev.createGeomType(GeomType::SPHERES);
    GeomType *cylType = dev.createGeomType(GeomType::CYLINDERS);

    Buffer *vb = dev.createBuffer(sizeof(vec3f) * nv, verts.data());
    const int half = nt / 2;
    Buffer *ib0 = dev.createBuffer(sizeof(vec3i) * half, idx.data());
    Buffer *ib1 = dev.createBuffer(sizeof(vec3i) * (nt - half), idx.data() + half);
    Buffer *cb0 = dev.createBuffer(sizeof(vec3f) * half, cols.data());
    Buffer *cb1 = dev.createBuffer(sizeof(vec3f) * (nt - half), cols.data() + half);
    Geom *t0 = triType->createGeom();
    t0->setVertices(vb, nv);
    t0->setIndices(ib0, half);
    t0->setPrimColors(cb0);
    t0->setUserID(0);
    Geom *t1 = triType->createGeom();
    t1->setVertices(vb, nv);
    t1->setIndices(ib1, nt - half);
    t1->setPrimColors(cb1);
    t1->setUserID(1);

    // spheres: uploaded in two parts to exercise Buffer::upload at an offset
    Buffer *sc = dev.createBuffer(sizeof(vec3f) * ns);
    const int h2 = ns / 2;
    sc->upload(centers.data(), sizeof(vec3f) * h2, 0);
    sc->upload(centers.data() + h2, sizeof(vec3f) * (ns - h2), sizeof(vec3f) * h2);
    Buffer *sr = dev.createBuffer(sizeof(float) * ns, radii.data());
    Geom *s0 = sphType->createGeom();
    s0->setSpheres(sc, sr, nullptr, ns, .01f);
    s0->setUserID(2);

    Buffer *cv = dev.createBuffer(sizeof(vec3f) * ncv, cverts.data());
    Buffer *ci = dev.createBuffer(sizeof(int32_t) * 2 * nc, cidx.data());
    Buffer *cr = dev.createBuffer(sizeof(float) * nc, cradii.data());
    Geom *c0 = cylType->createGeom();
    c0->setCylinders(cv, ncv, ci, nc, cr, nullptr, .01f);
    c0->setUserID(3);

    Group *gs = dev.createUserGeomsGroup({s0});
    Group *gt = dev.createTrianglesGroup({t0, t1});
    Group *gc = dev.createUserGeomsGroup({c0});
    gs->buildAccel();
    gt->buildAccel();
    gc->buildAccel();
    const affine3f I{{1, 0, 0}, {0, 1, 0}, {0, 0, 1}, {0, 0, 0}};
    Group *world = dev.createInstanceGroup({gs, gt, gc}, {0, 0, 0}, {I, I, I});
    world->buildAccel();

    // the reference's argument checks throw as BARNEY_RAISE does
    bool threw = false;
    try {
      Group *bad = dev.createInstanceGroup({gs, gt}, {0}, {I, I});
      dev.freeGroup(bad);
    } catch (const std::runtime_error &) {
      threw = true;
    }
    if (!threw) throw std::runtime_error("rtc_trace: mismatched createInstanceGroup did not throw");

    SetActiveGPU active(&dev);
    bnx_ray *d_rays = (bnx_ray *)dev.allocMem(sizeof(bnx_ray) * nr);
    bnx_hit *d_hits = (bnx_hit *)dev.allocMem(sizeof(bnx_hit) * nr);
    bnx_hit *h_hits = (bnx_hit *)dev.allocHost(3 * sizeof(bnx_hit) * nr);
    dev.copy(d_rays, rays.data(), sizeof(bnx_ray) * nr);
    auto trace_into = [&](Group *w, int k) {
      dev.memsetAsync(d_hits, 0xff, sizeof(bnx_hit) * nr);
      traceRays(w, d_rays, nr, d_hits);
      dev.copyAsync(h_hits + (size_t)k * nr, d_hits, sizeof(bnx_hit) * nr);
      dev.sync();
    };
    trace_into(world, 0);

    const affine3f M{{1, 0, 0}, {0, 1, 0}, {0, 0, 1}, move};
    world->setTransforms({I, M, I});
    world->refitAccel();
    trace_into(world, 1);

    // re-commit the triangles with every vertex moved
    std::vector<vec3f> moved(verts);
    for (vec3f &v : moved) v = vec3f{v.x + move.x, v.y + move.y, v.z + move.z};
    vb->upload(moved.data(), sizeof(vec3f) * nv);
    t0->setVertices(vb, nv);
    t0->setIndices(ib0, half);
    t1->setVertices(vb, nv);
    t1->setIndices(ib1, nt - half);
    Group *gt2 = dev.createTrianglesGroup({t0, t1});
    gt2->buildAccel();
    Group *world2 = dev.createInstanceGroup({gs, gt2, gc}, {0, 0, 0}, {I, I, I});
    world2->buildAccel();
    // the first groups still hold the geometries they were built over
    if (t0->retired.size() != 1 || t1->retired.size() != 1)
      throw std::runtime_error("rtc_trace: re-committed geometry did not keep the old one for its group");
    trace_into(world2, 2);
    dev.freeGroup(world);
    dev.freeGroup(gt);
    if (!t0->retired.empty() || !t1->retired.empty())
      throw std::runtime_error("rtc_trace: old geometry not freed with its last group");

    std::ofstream out(argv[2], std::ios::binary);
    out.write((const char *)h_hits, 3 * sizeof(bnx_hit) * nr);
    if (!out) throw std::runtime_error("rtc_trace: write failed");
    dev.freeHost(h_hits);
    dev.freeMem(d_hits);
    dev.freeMem(d_rays);

    for (Group *g : {world2, gt2, gs, gc}) dev.freeGroup(g);
    for (Geom *g : {t0, t1, s0, c0}) dev.freeGeom(g);
    for (Buffer *b : {vb, ib0, ib1, cb0, cb1, sc, sr, cv, ci, cr}) dev.freeBuffer(b);
    dev.freeGeomType(triType);
    dev.freeGeomType(sphType);
    dev.freeGeomType(cylType);
  } catch (const std::exception &e) {
    std::cerr << e.what() << std::endl;
    return 1;
  }
  std::printf("rtc_trace ok\n");
  return 0;
}
