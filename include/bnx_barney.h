/* bnx_barney.h -- barney's public API (native/include/barney.h:64-560,
 * ingowald/barney v1.4.0) over the MI355X backend.
 *
 * The same function names, handle types, enums and argument meaning as
 * barney.h, so barney's ANARI device (the barney/anari sources, which call only
 * these functions) and other bn* applications build against this header and
 * link libbnx.so instead of barney's native layer.  The functions are
 * extern "C" (barney.h declares them with C++ linkage inside BARNEY_NS):
 * plain pointers and sizes, callable from C, C++ and FFI alike.
 *
 * What is provided (each call cites the barney.h declaration it replaces):
 *   - one context on one GPU (bnContextCreate with one data group; the
 *     reference's multi-GPU / MPI data groups are out of scope);
 *   - data arrays, 2D / 3D texture data and 2D textures;
 *   - geometries "triangles", "spheres", "cylinders", "iso_surface"
 *     (native/geometry/Geometry.cpp:27-45);
 *   - materials "matte" / "AnariMatte" and "physicallyBased" / "AnariPBR"
 *     (material/HostMaterial.cpp:82-90), colour given as a value or as the
 *     name of a per-primitive attribute ("color" from bnSetString);
 *   - lights "directional" and "envmap" (light/Light.cpp:22-36), set on a
 *     group's "lights" data;
 *   - scalar fields "structured" and "unstructured" (tetrahedra) with
 *     volumes and bnVolumeSetXF;
 *   - cameras "perspective" (with depth of field), "orthographic", "omni";
 *   - the renderer ("bgColor", "ambientRadiance", "pathsPerPixel",
 *     "cutPlane"), frame buffers (float4, 8-bit RGBA and sRGB colour, depth,
 *     prim / inst / obj IDs) and bnRender.
 * Frame buffers denoise with a host-supplied denoiser (bnxContextSetDenoiser).
 * Unknown object types, parameters the backend does not use, and calls for
 * features out of scope (upscaling, MPI) are ignored with a warning on
 * stderr, as barney's warn_unsupported_object does; errors (bad handles,
 * failed builds) print the backend's message and leave the object unchanged.
 * bnGetLastError() returns the last message.
 *
 * Semantics kept from the reference:
 *   - parameters are staged by bnSet* and take effect at bnCommit;
 *   - bnBuild flattens a model slot's instances in the order of
 *     ModelSlot::flattenInstancesForDevice (native/ModelSlot.cpp:127-152):
 *     per instance its user geometries (spheres, then cylinders, one group
 *     per type here), then its triangles; the user instance ID of a hit is
 *     the instance's index in bnSetInstances;
 *   - bnRender adds renderer.pathsPerPixel samples to the frame, each
 *     bnRender accumulating until bnAccumReset (native/Context.cpp:139-182);
 *   - objects are reference counted: bnRelease drops the application's
 *     reference, objects that others still use stay alive.
 * Volumes and isosurfaces are marched in world space: their instance
 * transform must be the identity (a volume under another transform is
 * refused at bnBuild). */
#ifndef BNX_BARNEY_H
#define BNX_BARNEY_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
#define BNX_DEFAULT(v) = v
extern "C" {
#else
#define BNX_DEFAULT(v)
#endif

/* handles (barney.h:66-86) */
#ifdef __cplusplus
typedef struct _BNContext *BNContext;
typedef struct _BNObject {} *BNObject;
typedef struct _BNData : public _BNObject {} *BNData;
typedef struct _BNTextureData : public _BNObject {} *BNTextureData;
typedef struct _BNScalarField : public _BNObject {} *BNScalarField;
typedef struct _BNGeom : public _BNObject {} *BNGeom;
typedef struct _BNVolume : public _BNObject {} *BNVolume;
typedef struct _BNGroup : public _BNObject {} *BNGroup;
typedef struct _BNModel : public _BNObject {} *BNModel;
typedef struct _BNRenderer : public _BNObject {} *BNRenderer;
typedef struct _BNFrameBuffer : public _BNObject {} *BNFrameBuffer;
typedef struct _BNTexture2D : public _BNObject {} *BNTexture2D;
typedef struct _BNTexture3D : public _BNObject {} *BNTexture3D;
typedef struct _BNLight : public _BNObject {} *BNLight;
typedef struct _BNCamera : public _BNObject {} *BNCamera;
typedef struct _BNMaterial : public _BNObject {} *BNMaterial;
typedef struct _BNSampler : public _BNObject {} *BNSampler;
#else
typedef struct _BNContext *BNContext;
typedef struct _BNObject *BNObject;
typedef BNObject BNData, BNTextureData, BNScalarField, BNGeom, BNVolume, BNGroup, BNModel, BNRenderer,
    BNFrameBuffer, BNTexture2D, BNTexture3D, BNLight, BNCamera, BNMaterial, BNSampler;
#endif
typedef BNTexture2D BNTexture;

typedef struct bn_float2 { float x, y; } bn_float2;
typedef struct bn_float3 { float x, y, z; } bn_float3;
typedef struct bn_float4 { float x, y, z, w; } bn_float4;
typedef struct bn_int2 { int x, y; } bn_int2;
typedef struct bn_int3 { int x, y, z; } bn_int3;
typedef struct bn_int4 { int x, y, z, w; } bn_int4;

/* affine transform: linear part as columns vx, vy, vz, then p (barney.h:95-102) */
typedef struct BNTransform {
  struct {
    bn_float3 vx, vy, vz;
  } l;
  bn_float3 p;
} BNTransform;

/* barney.h:104-111 */
typedef enum {
  BN_FB_COLOR = (1 << 0),
  BN_FB_DEPTH = (1 << 1),
  BN_FB_PRIMID = (1 << 2),
  BN_FB_INSTID = (1 << 3),
  BN_FB_OBJID = (1 << 4),
  BN_FB_NORMAL = (1 << 5)
} BNFrameBufferChannel;

/* barney.h:113-197 (the values this backend reads) */
typedef enum {
  BN_DATA_UNDEFINED = 0,
  BN_DATA = 1,
  BN_OBJECT = 2,
  BN_TEXTURE = 3,
  BN_TEXTURE_3D = 4,
  BN_INT8 = 100,
  BN_UINT8 = 110,
  BN_INT32 = 140,
  BN_INT32_VEC2,
  BN_INT32_VEC3,
  BN_INT32_VEC4,
  BN_UINT32 = 150,
  BN_FLOAT32 = 180,
  BN_FLOAT32_VEC2,
  BN_FLOAT32_VEC3,
  BN_FLOAT32_VEC4,
  BN_INT = BN_INT32,
  BN_INT2 = BN_INT32_VEC2,
  BN_INT3 = BN_INT32_VEC3,
  BN_INT4 = BN_INT32_VEC4,
  BN_FLOAT = BN_FLOAT32,
  BN_FLOAT2 = BN_FLOAT32_VEC2,
  BN_FLOAT3 = BN_FLOAT32_VEC3,
  BN_FLOAT4 = BN_FLOAT32_VEC4,
  BN_UFIXED8 = 300,
  BN_UFIXED8_RGBA,
  BN_UFIXED8_RGBA_SRGB,
  BN_UFIXED16
} BNDataType;

/* barney.h:199-209 */
typedef enum {
  BN_UNSTRUCTURED_TET = 10,
  BN_UNSTRUCTURED_HEX = 12,
  BN_UNSTRUCTURED_PRISM = 13,
  BN_UNSTRUCTURED_PYRAMID = 14,
  BN_UNSTRUCTURED_POLYHEDRON = 42
} BNUnstructuredElementType;

typedef enum { BN_TEXTURE_NEAREST, BN_TEXTURE_LINEAR } BNTextureFilterMode;
typedef enum { BN_TEXTURE_WRAP, BN_TEXTURE_CLAMP, BN_TEXTURE_BORDER, BN_TEXTURE_MIRROR } BNTextureAddressMode;
typedef enum { BN_COLOR_SPACE_LINEAR, BN_COLOR_SPACE_SRGB } BNTextureColorSpace;

/* ------------------------------------------------------------ context */
/* barney.h:302-319: one data group on one GPU (gpuIDs[0], or GPU 0) */
BNContext bnContextCreate(int numDataGroupsOnThisContext BNX_DEFAULT(1),
                          const int *dataRanksInDataGroup BNX_DEFAULT(0), int numGPUs BNX_DEFAULT(-1),
                          const int *gpuIDs BNX_DEFAULT(0));
void bnContextDestroy(BNContext context);                      /* barney.h:325 */
int bnPhysicalDeviceCount(void);                               /* barney.h:263 */
void bnRelease(BNObject object);                               /* barney.h:335 */
void bnAddReference(BNObject object);                          /* barney.h:345 */
const char *bnGetLastError(void);                              /* (this backend) */
/* The denoiser of this context's frame buffers (this backend's extension:
 * barney's device creates its own, rtc::Device::createDenoiser, from OIDN
 * or OptiX).  A frame buffer with "enableDenoising" set runs it when its
 * colour channel is read, with "fadeOutDenoiser" (fb/FrameBuffer.cu:60-75,
 * 252-290); fn has bnx_frame_set_denoiser's signature and contract
 * (bnx_rt.h); NULL removes it. */
typedef int (*bnxDenoiseFn)(void *user, const float *d_in_rgba, const float *d_in_normal, float *d_out_rgba,
                            int32_t width, int32_t height, float blend, void *stream);
void bnxContextSetDenoiser(BNContext context, bnxDenoiseFn fn, void *user);

/* ---------------------------------------------------- object creation */
BNCamera bnCameraCreate(BNContext context, const char *type);  /* barney.h:237 */
BNFrameBuffer bnFrameBufferCreate(BNContext context, int deprecated BNX_DEFAULT(0)); /* :245 */
BNModel bnModelCreate(BNContext ctx);                          /* barney.h:248 */
BNRenderer bnRendererCreate(BNContext ctx, const char *type BNX_DEFAULT("default")); /* :253 */
BNData bnDataCreate(BNContext context, int whichSlot, BNDataType dataType, size_t numItems,
                    const void *items);                        /* barney.h:427 */
void bnDataSet(BNData data, size_t numItems, const void *items); /* barney.h:434 */
BNTextureData bnTextureData2DCreate(BNContext context, int whichSlot, BNDataType texelFormat, int width,
                                    int height, const void *items); /* barney.h:442 */
BNTextureData bnTextureData3DCreate(BNContext context, int whichSlot, BNDataType texelFormat, int width,
                                    int height, int depth, const void *items); /* barney.h:448 */
BNLight bnLightCreate(BNContext context, int whichSlot, const char *type); /* barney.h:455 */
BNGroup bnGroupCreate(BNContext context, int whichSlot, BNGeom *geoms, int numGeoms, BNVolume *volumes,
                      int numVolumes);                         /* barney.h:460 */
void bnGroupBuild(BNGroup group);                              /* barney.h:465 */
BNTexture2D bnTexture2DCreate(BNContext context, int whichSlot, BNDataType texelFormat, uint32_t size_x,
                              uint32_t size_y, const void *texels,
                              BNTextureFilterMode filterMode BNX_DEFAULT(BN_TEXTURE_LINEAR),
                              BNTextureAddressMode addressMode_x BNX_DEFAULT(BN_TEXTURE_WRAP),
                              BNTextureAddressMode addressMode_y BNX_DEFAULT(BN_TEXTURE_WRAP),
                              BNTextureColorSpace colorSpace BNX_DEFAULT(BN_COLOR_SPACE_LINEAR)); /* :468 */
BNGeom bnGeometryCreate(BNContext context, int whichSlot, const char *type); /* barney.h:504 */
BNScalarField bnScalarFieldCreate(BNContext context, int whichSlot, const char *type); /* barney.h:512 */
BNMaterial bnMaterialCreate(BNContext context, int whichSlot, const char *type); /* barney.h:519 */
BNSampler bnSamplerCreate(BNContext context, int whichSlot, const char *type); /* barney.h:524 */
BNVolume bnVolumeCreate(BNContext context, int whichSlot, BNScalarField sf); /* barney.h:530 */
void bnVolumeSetXF(BNVolume volume, bn_float2 domain, const bn_float4 *colorMap, int numColorMapValues,
                   float densityAt1);                          /* barney.h:535 */

/* ------------------------------------------------------- parameters */
void bnCommit(BNObject target);                                /* barney.h:266 */
void bnSetString(BNObject target, const char *paramName, const char *value); /* :269 */
void bnSetData(BNObject target, const char *paramName, BNData value); /* :272 */
void bnSetObject(BNObject target, const char *paramName, BNObject value); /* :275 */
void bnSetLight(BNObject target, const char *paramName, BNLight value); /* :278 */
void bnSet1i(BNObject target, const char *paramName, int value);
void bnSet2i(BNObject target, const char *paramName, int x, int y);
void bnSet3i(BNObject target, const char *paramName, int x, int y, int z);
void bnSet4i(BNObject target, const char *paramName, int x, int y, int z, int w);
void bnSet1f(BNObject target, const char *paramName, float value);
void bnSet2f(BNObject target, const char *paramName, float x, float y);
void bnSet3f(BNObject target, const char *paramName, float x, float y, float z);
void bnSet4f(BNObject target, const char *paramName, float x, float y, float z, float w);
void bnSet4x3fv(BNObject target, const char *paramName, const BNTransform *affineMatrix); /* :307 */
void bnSet4x4fv(BNObject target, const char *paramName, const bn_float4 *xfm); /* :310 */

/* ------------------------------------------------------- model, render */
void bnSetInstances(BNModel model, int whichSlot, BNGroup *groupsToInstantiate, BNTransform *instanceTransforms,
                    int numInstances);                         /* barney.h:392 */
void bnUpdateInstanceTransforms(BNModel model, int whichSlot, BNTransform *instanceTransforms,
                                int numInstances);             /* barney.h:403 */
void bnSetInstanceAttributes(BNModel model, int whichSlot, const char *attributeName, BNData data); /* :411 */
void bnBuild(BNModel model, int whichDataSlot);                /* barney.h:354 */
void bnAccumReset(BNFrameBuffer fb);                           /* barney.h:365 */
void bnFrameBufferResize(BNFrameBuffer fb, BNDataType colorFormat, int sizeX, int sizeY,
                         uint32_t requiredChannels BNX_DEFAULT(BN_FB_COLOR)); /* barney.h:368 */
void bnFrameBufferRead(BNFrameBuffer fb, BNFrameBufferChannel channelToRead, void *pointerToReadDataInto,
                       BNDataType requiredFormat);             /* barney.h:374 */
void bnFrameBufferGetSize(BNFrameBuffer fb, int *sizeX, int *sizeY); /* barney.h:383 */
void bnRender(BNRenderer renderer, BNModel model, BNCamera camera, BNFrameBuffer fb); /* barney.h:386 */

#ifdef __cplusplus
} /* extern "C" */
static_assert(sizeof(BNTransform) == 48, "BNTransform layout");
#endif
#undef BNX_DEFAULT

#endif /* BNX_BARNEY_H */
