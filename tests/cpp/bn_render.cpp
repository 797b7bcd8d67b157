// bn_render.cpp -- a frame rendered through barney's public bn* API
// (include/bnx_barney.h, after native/include/barney.h), as barney's ANARI
// device drives it: bnContextCreate, bnDataCreate, bnGeometryCreate +
// bnSetData/bnSetString/bnSetObject/bnCommit, bnMaterialCreate,
// bnLightCreate, bnGroupCreate + "lights" + bnGroupBuild, bnModelCreate +
// bnSetInstances + bnBuild, bnCameraCreate, bnRendererCreate,
// bnFrameBufferCreate + bnFrameBufferResize, bnRender, bnFrameBufferRead,
// bnRelease, bnContextDestroy.
//
// usage: bn_render <in> <out>
//   in:  int32 mode (0: surfaces, 1: structured volume), int32 W, H, spp,
//        float3 camera position, direction, up, float fovy (degrees),
//        float4 background, float ambient, float3 light direction (the
//        emission direction), float radiance, float3 light colour; then
//        mode 0: int32 nv, float3[nv], int32 nt, int3[nt], float3[nt]
//                triangle colours, int32 ns, float3[ns] centres, float[ns]
//                radii, float3[ns] sphere colours (triangles user ID 0,
//                spheres user ID 1, one group, matte with colour from the
//                per-primitive "color")
//        mode 1: int32 dims[3], float3 origin, float3 spacing, float[n]
//                scalars, int32 ntf, float4[ntf], float2 domain, float
//                density
//   out: float4 colour[W H], int32 primID, instID, objID [W H], float
//        depth [W H], uint32 sRGB8 colour [W H] (rows from the bottom),
//        float4 colour [W H] read again with "enableDenoising" set and a
//        context denoiser (bnxContextSetDenoiser) that writes 1 - rgb and
//        keeps alpha, then float4 colour [W H] with denoising off again
// Prints "bn_render ok" and exits 0 on success.
#include "bnx_barney.h"
#include "bnx_rt.h"

#include <cstdio>
#include <fstream>
#include <iostream>
#include <stdexcept>
#include <string>
#include <vector>

template <typename T> static std::vector<T> rd(std::ifstream &f, size_t n)
{
  std::vector<T> v(n);
  f.read((char *)v.data(), sizeof(T) * n);
  if (!f) throw std::runtime_error("bn_render: short input");
  return v;
}
template <typename T> static T rd1(std::ifstream &f) { return rd<T>(f, 1)[0]; }

// a denoiser that writes 1 - rgb and keeps alpha (bnx_copy without a
// device handle: the bn* API exposes none)
static int invert_rgb(void *, const float *in, const float *, float *out, int32_t w, int32_t h, float, void *)
{
  std::vector<float> v(4 * (size_t)w * h);
  if (bnx_copy(nullptr, v.data(), in, sizeof(float) * v.size()) != 0) return 1;
  for (size_t i = 0; i < v.size(); i++)
    if (i % 4 != 3) v[i] = 1.f - v[i];
  return bnx_copy(nullptr, out, v.data(), sizeof(float) * v.size()) != 0 ? 2 : 0;
}

static void ok_or_throw(const void *h, const char *what)
{
  if (!h) throw std::runtime_error(std::string(what) + " failed: " + bnGetLastError());
}

int main(int argc, char **argv)
{
  if (argc != 3) {
    std::cerr << "usage: bn_render <in> <out>" << std::endl;
    return 2;
  }
  try {
    std::ifstream in(argv[1], std::ios::binary);
    const int mode = rd1<int32_t>(in), W = rd1<int32_t>(in), H = rd1<int32_t>(in), spp = rd1<int32_t>(in);
    auto cam = rd<float>(in, 10);
    auto bg = rd<float>(in, 4);
    const float ambient = rd1<float>(in);
    auto light = rd<float>(in, 7);

    if (bnPhysicalDeviceCount() < 1) throw std::runtime_error("bn_render: no GPU");
    const int gpu = 0;
    BNContext ctx = bnContextCreate(1, nullptr, 1, &gpu);
    ok_or_throw(ctx, "bnContextCreate");
    std::vector<BNGeom> geoms;
    std::vector<BNVolume> volumes;

    if (mode == 0) {
      const int nv = rd1<int32_t>(in);
      auto verts = rd<float>(in, 3 * (size_t)nv);
      const int nt = rd1<int32_t>(in);
      auto idx = rd<int32_t>(in, 3 * (size_t)nt);
      auto tcol = rd<float>(in, 3 * (size_t)nt);
      const int ns = rd1<int32_t>(in);
      auto centers = rd<float>(in, 3 * (size_t)ns);
      auto radii = rd<float>(in, ns);
      auto scol = rd<float>(in, 3 * (size_t)ns);

      BNMaterial mat = bnMaterialCreate(ctx, 0, "matte");
      ok_or_throw(mat, "bnMaterialCreate");
      bnSetString(mat, "color", "color");
      bnCommit(mat);

      BNGeom tri = bnGeometryCreate(ctx, 0, "triangles");
      BNData v = bnDataCreate(ctx, 0, BN_FLOAT3, nv, verts.data());
      BNData ix = bnDataCreate(ctx, 0, BN_INT3, nt, idx.data());
      BNData tc = bnDataCreate(ctx, 0, BN_FLOAT3, nt, tcol.data());
      bnSetData(tri, "vertices", v);
      bnSetData(tri, "indices", ix);
      bnSetData(tri, "primitive.color", tc);
      bnSetObject(tri, "material", mat);
      bnSet1i(tri, "userID", 0);
      bnCommit(tri);

      BNGeom sph = bnGeometryCreate(ctx, 0, "spheres");
      BNData c = bnDataCreate(ctx, 0, BN_FLOAT3, ns, centers.data());
      BNData r = bnDataCreate(ctx, 0, BN_FLOAT, ns, radii.data());
      BNData sc = bnDataCreate(ctx, 0, BN_FLOAT3, ns, scol.data());
      bnSetData(sph, "origins", c);
      bnSetData(sph, "radii", r);
      bnSetData(sph, "primitive.color", sc);
      bnSetObject(sph, "material", mat);
      bnSet1i(sph, "userID", 1);
      bnCommit(sph);
      // the geometries hold the data and the material now
      for (BNObject o : {(BNObject)v, (BNObject)ix, (BNObject)tc, (BNObject)c, (BNObject)r, (BNObject)sc,
                         (BNObject)mat})
        bnRelease(o);
      geoms = {sph, tri};
    } else {
      auto dims = rd<int32_t>(in, 3);
      auto origin = rd<float>(in, 3);
      auto spacing = rd<float>(in, 3);
      auto scalars = rd<float>(in, (size_t)dims[0] * dims[1] * dims[2]);
      const int ntf = rd1<int32_t>(in);
      auto tf = rd<bn_float4>(in, ntf);
      auto dom = rd<float>(in, 2);
      const float density = rd1<float>(in);
      BNTextureData td = bnTextureData3DCreate(ctx, 0, BN_FLOAT, dims[0], dims[1], dims[2], scalars.data());
      BNScalarField sf = bnScalarFieldCreate(ctx, 0, "structured");
      bnSetObject(sf, "textureData", td);
      bnSet3i(sf, "dims", dims[0], dims[1], dims[2]);
      bnSet3f(sf, "gridOrigin", origin[0], origin[1], origin[2]);
      bnSet3f(sf, "gridSpacing", spacing[0], spacing[1], spacing[2]);
      bnCommit(sf);
      BNVolume vol = bnVolumeCreate(ctx, 0, sf);
      ok_or_throw(vol, "bnVolumeCreate");
      bnVolumeSetXF(vol, bn_float2{dom[0], dom[1]}, tf.data(), ntf, density);
      bnCommit(vol);
      bnRelease(td);
      bnRelease(sf);
      volumes = {vol};
    }

    BNLight dl = bnLightCreate(ctx, 0, "directional");
    bnSet3f(dl, "direction", light[0], light[1], light[2]);
    bnSet1f(dl, "radiance", light[3]);
    bnSet3f(dl, "color", light[4], light[5], light[6]);
    bnCommit(dl);
    BNData lights = bnDataCreate(ctx, 0, BN_OBJECT, 1, &dl);

    BNGroup group = bnGroupCreate(ctx, 0, geoms.data(), (int)geoms.size(), volumes.data(), (int)volumes.size());
    ok_or_throw(group, "bnGroupCreate");
    bnSetData(group, "lights", lights);
    bnCommit(group);
    bnGroupBuild(group);

    BNModel model = bnModelCreate(ctx);
    BNTransform I{};
    I.l.vx.x = I.l.vy.y = I.l.vz.z = 1.f;
    bnSetInstances(model, 0, &group, &I, 1);
    bnBuild(model, 0);

    BNCamera camera = bnCameraCreate(ctx, "perspective");
    bnSet3f(camera, "position", cam[0], cam[1], cam[2]);
    bnSet3f(camera, "direction", cam[3], cam[4], cam[5]);
    bnSet3f(camera, "up", cam[6], cam[7], cam[8]);
    bnSet1f(camera, "fovy", cam[9]);
    bnSet1f(camera, "aspect", (float)W / (float)H);
    bnCommit(camera);

    BNRenderer renderer = bnRendererCreate(ctx, "default");
    bnSet4f(renderer, "bgColor", bg[0], bg[1], bg[2], bg[3]);
    bnSet1f(renderer, "ambientRadiance", ambient);
    bnSet1i(renderer, "pathsPerPixel", spp);
    bnCommit(renderer);

    BNFrameBuffer fb = bnFrameBufferCreate(ctx, 0);
    bnFrameBufferResize(fb, BN_FLOAT4, W, H, BN_FB_COLOR | BN_FB_DEPTH | BN_FB_PRIMID | BN_FB_INSTID | BN_FB_OBJID);
    int sx = 0, sy = 0;
    bnFrameBufferGetSize(fb, &sx, &sy);
    if (sx != W || sy != H) throw std::runtime_error("bn_render: frame size differs");
    bnRender(renderer, model, camera, fb);

    const size_t np = (size_t)W * H;
    std::vector<float> color(4 * np), depth(np);
    std::vector<int32_t> prim(np), inst(np), obj(np);
    std::vector<uint32_t> srgb(np);
    bnFrameBufferRead(fb, BN_FB_COLOR, color.data(), BN_FLOAT4);
    bnFrameBufferRead(fb, BN_FB_PRIMID, prim.data(), BN_INT);
    bnFrameBufferRead(fb, BN_FB_INSTID, inst.data(), BN_INT);
    bnFrameBufferRead(fb, BN_FB_OBJID, obj.data(), BN_INT);
    bnFrameBufferRead(fb, BN_FB_DEPTH, depth.data(), BN_FLOAT);
    bnFrameBufferRead(fb, BN_FB_COLOR, srgb.data(), BN_UFIXED8_RGBA_SRGB);
    // denoising with a context denoiser, then off again
    std::vector<float> denoised(4 * np), plain_again(4 * np);
    bnxContextSetDenoiser(ctx, invert_rgb, nullptr);
    bnSet1i(fb, "enableDenoising", 1);
    bnFrameBufferRead(fb, BN_FB_COLOR, denoised.data(), BN_FLOAT4);
    bnSet1i(fb, "enableDenoising", 0);
    bnFrameBufferRead(fb, BN_FB_COLOR, plain_again.data(), BN_FLOAT4);
    if (std::string(bnGetLastError()).find("failed") != std::string::npos)
      throw std::runtime_error(std::string("bn_render: ") + bnGetLastError());

    std::ofstream out(argv[2], std::ios::binary);
    out.write((const char *)color.data(), sizeof(float) * color.size());
    out.write((const char *)prim.data(), 4 * np);
    out.write((const char *)inst.data(), 4 * np);
    out.write((const char *)obj.data(), 4 * np);
    out.write((const char *)depth.data(), 4 * np);
    out.write((const char *)srgb.data(), 4 * np);
    out.write((const char *)denoised.data(), sizeof(float) * denoised.size());
    out.write((const char *)plain_again.data(), sizeof(float) * plain_again.size());
    if (!out) throw std::runtime_error("bn_render: write failed");

    for (BNObject o : {(BNObject)fb, (BNObject)renderer, (BNObject)camera, (BNObject)model, (BNObject)group,
                       (BNObject)lights, (BNObject)dl})
      bnRelease(o);
    for (BNGeom g : geoms) bnRelease(g);
    for (BNVolume v : volumes) bnRelease(v);
    bnContextDestroy(ctx);
  } catch (const std::exception &e) {
    std::cerr << e.what() << std::endl;
    return 1;
  }
  std::printf("bn_render ok\n");
  return 0;
}
