// DeviceHip: ProtoRay's Device (render/device.h:56-114) for the stream path renderer
// ("pathStream") on one MI355X, forwarding every call to the C ABI of libprt_hip.so
// (include/prt_hip.h).  A ProtoRay maintainer adds this file and device_hip.cpp to the protoray
// sources and selects the device where main/render/render.cpp:233-241 picks DeviceCpu.
//
// Inputs are read as DeviceCpu reads them (render/device_cpu.cpp), with these differences:
// * initScene loads .rkscene / .gltf / .mesh files with the library's own loaders
//   (prt_scene_file_*; instances flattened) instead of the reference Scene;
// * image lights and the sun/sky light need a host-side environment (the lat-long image, or the
//   baked sun + sky pair); DeviceHip takes it from an EnvironmentProvider the caller installs
//   (in ProtoRay: an adapter over loadImageTexture and SunSkyLight), ambient and distant lights
//   are read from the props directly;
// * AOVs other than colour (with variance and half buffer), albedo, albedo1, normal, normal1 and
//   depth, the pixel filters' splat mode and motion vectors (the previous / next cameras) are not
//   supported and are rejected with std::invalid_argument.
#pragma once

#include <functional>
#include <memory>
#include <string>
#include <vector>

#include "render/device.h"
#include "image/tone_mapper.h"
#include "prt_hip.h"

namespace prt {

// The environment of initLights for the props the device does not read itself ("imageLight",
// "sunSkyLight*"): lights (image lights of kind PRT_LIGHT_IMAGE refer to rgb) and one lat-long
// image (texture row order, row 0 = nadir).  Returns false when the props name no such light.
struct HipEnvironment
{
  std::vector<prt_light> lights;
  std::vector<float> rgb;
  int width = 0, height = 0;
  int lutSize = 0;
};
using EnvironmentProvider = std::function<bool(const Props& props, float sceneRadius, HipEnvironment& env)>;

class DeviceHip : public Device
{
public:
  explicit DeviceHip(int deviceOrdinal = 0);
  ~DeviceHip() override;
  DeviceHip(const DeviceHip&) = delete;
  DeviceHip& operator =(const DeviceHip&) = delete;

  void setEnvironmentProvider(EnvironmentProvider provider) { environment = std::move(provider); }
  prt_context* getContext() const { return ctx; }

  std::string getInfo() override;

  // Scene
  void initScene(const std::string& path, const Props& props) override;
  Box3f getSceneBounds() override;
  void initLights(const Props& props) override;

  // Renderer
  void initRenderer(const Props& props, Props& stats) override;
  void initSampler(unsigned int seed) override;
  void render(Props& stats) override;
  Props queryPixel(const Vec2f& point) override;
  void queryPixel(const Vec2f& point, Array<Props>& hits) override;
  Props queryPixel(int x, int y) override;
  void queryPixel(int x, int y, Array<Props>& hits) override;
  Props queryRay(const Ray& ray) override;
  bool queryOcclusionRay(const Ray& ray) override;

  // Camera
  void initCamera(const Props& props) override;
  void initCamera(const Props& props, const Props& prevProps, const Props& nextProps) override;
  Mat4f getWorldToViewD3D() override;
  Mat4f getViewToClipD3D() override;
  Mat4f getWorldToRaster() override;

  // Jitter
  bool hasJitter() override;
  Vec2f getJitter() override;
  void setJitter(const Vec2f& jitter) override;
  void disableJitter() override;

  // Framebuffer
  void initFrame(const Vec2i& size, const Props& props) override;
  void initFilter(const Props& props) override;
  void initToneMapper(const Props& props) override;
  void clearFrame() override;
  void blitFrame(Surface& dest) override;
  void blitFrame(FrameAovs& dest) override;
  void blitFrameLdr(Surface& dest, const float* src, float offset = 0.f, float scale = 1.f) override;
  void blitFrameLdr(Surface& dest, const Vec3f* src, float offset = 0.f, float scale = 1.f) override;
  void blitFrameHdr(Surface& dest, const Vec3f* src) override;
  float getFrameError() override;
  float getFilterWidth() const override;

private:
  void matrices(float* worldToView, float* viewToClip, float* worldToRaster);
  Props hitProps(const prt_pixel_hit& h, bool withDepth) const;

  prt_context* ctx = nullptr;
  EnvironmentProvider environment;
  std::vector<std::string> materialNames;
  prt_camera camera{};
  Vec2i cameraRaster = Vec2i(0);
  bool hasCamera = false;
  Vec2i frameSize = Vec2i(0);
  bool halfBuffer = false;
  std::shared_ptr<ToneMapper> toneMapper;
};

} // namespace prt
