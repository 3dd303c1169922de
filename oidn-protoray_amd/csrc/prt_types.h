// Plain-old-data records shared by the host driver and the gfx950 kernels.
#pragma once

#include <stdint.h>

#define PRT_KIND_DIFFUSE 0
#define PRT_KIND_CONDUCTOR 1
#define PRT_KIND_DIELECTRIC 2
#define PRT_KIND_COATED 3
#define PRT_KIND_THIN_DIELECTRIC 4
#define PRT_KIND_STANDARD 5
// the transparent BTDF as a base of a kind-5 material's mix (a constants record, never a material kind)
#define PRT_KIND_TRANSPARENT 6

// Bins of the guide tables that narrow the searches of the environment's sampling CDFs (a power
// of two, so s * PRT_ENV_GUIDE is exact in float)
#define PRT_ENV_GUIDE 32

#define PRT_LIGHT_AMBIENT 0
#define PRT_LIGHT_DISTANT 1
#define PRT_LIGHT_IMAGE 2
#define PRT_MAX_LIGHTS 8

// Energy-compensation LUTs, one flat float array (bsdf/albedo_tables.h:14-157):
//   MicrofacetAlbedoTable dir [32 roughness][32 cos], avg [32]
//   MicrofacetDielectricAlbedoTable etaDir [16 rough][16 eta][16 cos] (eta in [1/3,1]), etaAvg
//   [16][16], rcpEtaDir (eta in [1,3]), rcpEtaAvg
//   MicrofacetDielectricReflectionAlbedoTable etaDir [16 rough][16 eta][16 cos] (eta in
//   [1/3,1]), etaAvg [16][16] (bsdf/albedo_tables.h:160-205, read by the coating)
#define PRT_LUT_GGX_DIR 0
#define PRT_LUT_GGX_AVG 1024
#define PRT_LUT_DIEL_ETA_DIR 1056
#define PRT_LUT_DIEL_ETA_AVG 5152
#define PRT_LUT_DIEL_RCPETA_DIR 5408
#define PRT_LUT_DIEL_RCPETA_AVG 9504
#define PRT_LUT_DIEL_REFL_DIR 9760
#define PRT_LUT_DIEL_REFL_AVG 13856
//   MicrofacetSheenAlbedoTable dir [16 roughness][16 cos] (bsdf/albedo_tables.h:281-300)
#define PRT_LUT_SHEEN_DIR 14112
#define PRT_LUT_FLOATS 14368

// Per-material constants derived on the host at upload (the per-hit constructor work of
// RoughConductorBrdfSimd, rough_conductor.h:121-130, RoughDielectricBsdfSimd,
// rough_dielectric.h:279-291, and ThinRoughDielectricBsdfSimd, thin_rough_dielectric.h:180-192,
// depends on the material only).  312 bytes.
struct prt_material_constants
{
  int32_t kind;
  float roughness;   // clamped specularRoughness
  float alpha;       // max(roughness^2, 0.001)
  float importance;  // MultiBsdf importance of the single lobe (diffuse: max(color))
  float c0[3];       // diffuse color | conductor eta | coated: diffuse color | thin: color
  float c1[3];       // conductor k | coated: coat color (transmittance) | thin: fmsScale in x
  float c2[3];       // conductor fmsScale | coated: thickness, weight (scale), diffuse importance
                     // | thin: thickness in x
  float rr[3];       // conductor reflectivity | coated, thin: albedo (the albedo AOV value)
  float Eavg;        // conductor, thin: GGX average albedo | coated: Ravg (average coat
                     // reflectance x scale)
  float eta[2];      // dielectric eta seen from outside [0] / inside [1] | coated, thin: eta (<= 1)
  float EavgEta[2];
  float EavgRcpEta[2];
  float fmsRatio[2];
  int32_t mediumInt;  // interior medium ID (0 for non-dielectrics)
  float emission[3];  // emissionColor * emissionLuminance
  int32_t emissive;   // hits end the path as area-light hits (render/path_renderer_stream.h:419-582)
  // read by the adjoint's parameter derivatives only
  float edge[3];       // conductor edge tint g
  float Favg[3];       // conductor Kulla-Conty average Fresnel
  float FavgEta[2];    // dielectric average Fresnel at eta, outside [0] / inside [1]
  float FavgRcpEta[2]; // dielectric average Fresnel at 1 / eta
  // opacity mix (material/standard.cpp:412-420, :613-627): mix 0 = the lobe alone, 1 = the lobe
  // (scale opacity) and a transparent BTDF (scale transp) in a MultiBsdf, 2 = transparent only
  int32_t mix;
  float opacity, transp;
  int32_t transparent;     // shadow rays pass through, attenuated (render/transparent_filter.h)
  float diffuseRoughness;  // kinds 0 and 3: Oren-Nayar base when > 0 (bsdf/rough_diffuse.h)
  // sheen layer over the lobe (material/standard.cpp:599-611, bsdf/sheen.h:199-451)
  int32_t sheen;           // sheen weight above 1e-5
  float sheenColor[3];
  float sheenScale, sheenRoughness, sheenAlpha;
  int32_t transmissive;    // MaterialType::Transmissive (material/standard.cpp:159-160)
  int32_t mediaInterface;  // a transmissive surface that is not thin-walled (material/standard.cpp:159-175)
  int32_t smooth;          // a diffuse or glossy lobe: next-event estimation applies (path_renderer_stream.h:659)
  // kind 5, StandardMaterial's bases (material/standard.cpp:409-566): the base records seen from
  // outside [0] and inside [1] (consecutive records at this record's index + kidBase[o]), their scales and
  // importances (base importance x scale), the mix's importance, whether every base is diffuse,
  // and the coat over the mix (eta folded to <= 1 and Ravg per orientation); rr holds the albedo
  int32_t nKids;
  int32_t kidBase[2];
  float kidScale[3], kidImp[3];
  float impSum;
  int32_t mixDiffuseOnly;
  int32_t coat;
  float coatEta[2], coatRavg[2];
  float coatColor[3];
  float coatThickness, coatScale, coatRoughness, coatAlpha;
  // textured baseColor of kinds 0 and 3 (material/material_param.h:37-41, material/standard.cpp:
  // 424): 1 + texture index (0: none) and the parameter's scale, the material's (r, g, b)
  int32_t texBase;
  float texScale[3];
};
static_assert(sizeof(prt_material_constants) == 328, "material constants: 328 bytes");

// Image texture as the kernels read it (image/image_texture.h:148-227): texels decoded to linear
// float4 at upload, rows bottom to top (the reference stores images upside down,
// image_texture.h:352-353), from texel `offset` of the scene's texel array.  48 bytes.
#define PRT_TEX_ADDR_REPEAT 0
#define PRT_TEX_ADDR_CLAMP 1
struct prt_texture_dev
{
  int32_t offset;
  int32_t w, h;
  int32_t addrU, addrV;
  float scaleU, scaleV, offU, offV;  // uv * scale + offset (image_texture.h:266-269)
  float halfU, halfV;                // half-texel size 0.5 / size (:163)
  int32_t pad;
};

// An instance of a mesh (geometry/instance.h:12-60) as the kernels read it: the world -> object
// affine map that takes rays into the mesh's space (Embree's instance traversal), the object ->
// world linear part that takes the frame's tangents back (xfmVector) and its inverse transpose
// that takes the normals back (xfmNormal), all as rows.  128 bytes.
struct prt_instance_dev
{
  float inv[12];  // x_obj = inv[0..2] . x + inv[3], ... (rows of [A | b])
  float lin[9];   // rows of L, x_world = L x_obj + t
  float nrm[9];   // rows of L^-T
  int32_t mesh;
  int32_t pad[1];
};

// Environment light record as the kernels read it.  64 bytes.
struct prt_light_dev
{
  int32_t kind;
  float r, g, b;
  float dx, dy, dz;
  float cosAngle;  // distant: cos(half angle)
  float conePdf;   // distant: 1 / (4 pi sin^2(angle/2)) (sampling/shape_sampler.h:100-123)
  float fu[3];     // distant: sampling frame U (math/basis3.h:173-180)
  float fv[3];     // distant: sampling frame V
  float pad;
};

// 4-wide BVH node (128 bytes, one L2 line): the boxes of up to four children as SoA (lo/hi per
// axis for children 0..3) and four child references: child >= 0 is an inner node index,
// child < 0 a leaf with (child & 0x7fffffff) = first_triangle << 3 | (count - 1), -1 an unused
// slot (which also holds an empty box, lo = +inf, hi = -inf).
struct prt_bvh4_node
{
  float lox[4], hix[4];
  float loy[4], hiy[4];
  float loz[4], hiz[4];
  int32_t child[4];
  int32_t pad[4];
};

// Triangles per BVH leaf at most: the builder's cost model may make leaves up to PRT_LEAF_MAX;
// a leaf reference encodes up to PRT_LEAF_CAP.
#ifndef PRT_LEAF_MAX
#define PRT_LEAF_MAX 2
#endif
#define PRT_LEAF_CAP 8
// Traversal stack entries per lane kept in LDS (8 B each: node + entry distance) and resident
// blocks of 256 lanes per CU for the persistent traversal kernels.  11 x 8 B x 256 = 22 KB per
// block, 7 blocks per CU (154 KB of the 160 KB LDS, 7 waves per SIMD): measured best on the
// 1080p workload against (10, 7), (16, 4), (12, 5), (12, 6), (9, 8) and (8, 8).
#ifndef PRT_LDS_STACK
#define PRT_LDS_STACK 11
#endif
// The shadow-ray kernel's stacks hold nodes only (any-hit traversal never reads the entry
// distances): 24 x 4 B x 256 = 24 KB per block at its 6 blocks per CU.
#ifndef PRT_LDS_STACK_ANYHIT
#define PRT_LDS_STACK_ANYHIT 24
#endif
#define PRT_MAX_STACK 128
// Path vertices the adjoint records per lane and replay (deeper paths are replayed once more per
// further group of this many depths).
#define PRT_ADJ_MAX_VERTICES 8
#ifndef PRT_TRACE_BLOCKS_PER_CU
#define PRT_TRACE_BLOCKS_PER_CU 7
#endif
// Blocks per CU of the opaque-scene shadow-ray kernel: one fewer than the closest-hit kernel's
// seven (measured on the 1080p workload: 5.53 against 5.62 ms of shadow rays per sample pass;
// closest hit is fastest at seven, 7.46 ms against 7.54 at six).
#ifndef PRT_OCCLUDED_BLOCKS_PER_CU
#define PRT_OCCLUDED_BLOCKS_PER_CU 6
#endif

// Path flags (per pixel, one int).
#define PRT_F_LASTTYPE_MASK 0xff
#define PRT_F_AOV1 (1 << 8)       // first-hit AOVs written  (path_renderer_stream.h:75)
#define PRT_F_AOV2 (1 << 9)       // other AOVs written       (path_renderer_stream.h:76)
#define PRT_F_DIRECT (1 << 10)    // shadow ray pending       (path_renderer_stream.h:74)
#define PRT_F_TERMINATE (1 << 11) // path ends after the shadow test
#define PRT_F_MEDIUM_SHIFT 16

struct prt_frame_dev  // device pointers of the frame accumulators (FP32 sums, planar)
{
  float* color;   // 3 x npix
  float* weight;  // npix
  float* color2;  // 3 x npix (sum of squares) or null
  float* colorB;  // 3 x npix: sums of the class-B samples only (the half buffer) or null
  float* weightB; // npix
  float* albedo; float* albedoW;
  float* albedo1; float* albedo1W;
  float* normal; float* normalW;
  float* normal1; float* normal1W;
  float* depth; float* depthW;
  float* dbgL;          // 3 x npix or null
  uint32_t* dbgHash;    // npix or null
  int32_t* dbgFirst;    // npix or null
};
