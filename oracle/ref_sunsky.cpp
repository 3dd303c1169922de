// Sun/sky bake for scene inputs: TEST/INPUT INFRASTRUCTURE, run only in the build container.
//
// light/sun_sky_light.cpp cannot be compiled here (it includes sys/tasking.h, whose TBB
// headers are absent), so this program restates its bake loop (sun_sky_light.cpp:17-193,
// with the default hue/saturation/scale settings) on top of the reference's own vendored
// Hosek-Wilkie spectral sky model (light/sky_model/sky_model.cpp, BSD-3) and CIE tables
// (image/color.cpp), compiled from /root/reference by oracle/Makefile.ref.
//
// Output (raw little-endian):
//   <out>/sky.bin   float32 [256][512][3], row y = polar angle (y+0.5)/256*pi from the zenith
//   <out>/sun.txt   "r g b radius dx dy dz": sun radiance, angular radius (rad), direction

#include <cstdio>
#include <vector>
#include "math/integration.h"
#include "math/basis3.h"
#include "sampling/shape_sampler.h"
#include "image/color.h"
#include "light/sky_model/sky_model.h"

using namespace prt;

int main(int argc, char** argv)
{
  if (argc < 7) { fprintf(stderr, "usage: %s outdir dx dy dz turbidity albedo\n", argv[0]); return 2; }
  const std::string out = argv[1];
  Vec3f sunDir = normalize(Vec3f(std::stof(argv[2]), std::stof(argv[3]), std::stof(argv[4])));
  const float turbidity = std::stof(argv[5]);
  const float albedo = std::stof(argv[6]);
  const float lambdaMin = 320.0f, lambdaMax = 720.0f;
  const float horizon = 0.01f; // clamp(default 0, 0.01, 1)

  double sunTheta = std::min(double(acosSafe(sunDir.y)), 0.999 * double(pi) / 2.0);
  double sunPhi = std::atan2(sunDir.x, -sunDir.z);
  double sunElevation = double(pi) / 2.0 - sunTheta;
  ArHosekSkyModelState* model = arhosekskymodelstate_alloc_init(sunElevation, turbidity, albedo);
  float sunRadius = float(model->solar_radius);

  auto sunRadianceFunc = [=](Vec2f s)
  {
    float pdf;
    s = min(s, Vec2f(1.0f - 1e-6f));
    Vec3f df = makeFrame(sunDir) * uniformSampleCone(pdf, s, sunRadius);
    Vec3d d(df.x, df.y, df.z);
    double theta = std::min(double(acosSafe(d.y)), 0.999 * double(pi) / 2.0);
    double phi = std::atan2(d.x, -d.z);
    double cosGamma = std::cos(theta) * std::cos(sunTheta) + std::sin(theta) * std::sin(sunTheta) * std::cos(phi - sunPhi);
    double gamma = acosSafe(cosGamma);
    Vec3f radiance = zero;
    float radianceNorm = zero;
    for (int i = 0; i < cieSize; ++i)
      if (cieLambda[i] >= lambdaMin && cieLambda[i] <= lambdaMax)
      {
        double r = arhosekskymodel_solar_radiance(model, theta, gamma, cieLambda[i]) - arhosekskymodel_radiance(model, theta, gamma, cieLambda[i]);
        radiance += float(r) * cieXyz(i);
        radianceNorm += cieY[i];
      }
    return xyzToRgb(radiance / radianceNorm);
  };
  Vec3f sunRadiance = integrate2D(sunRadianceFunc, Vec2f(0.0f), Vec2f(1.0f), 5.0f);

  const int W = 512, H = 256;
  std::vector<float> sky(size_t(W) * H * 3);
  for (int y = 0; y < H; ++y)
    for (int x = 0; x < W; ++x)
    {
      const double maxTheta = 0.999 * double(pi) / 2.0;
      const double maxThetaHorizon = (horizon + 1.0) * double(pi) / 2.0;
      Vec3f radiance = zero;
      double theta = (double(y) + 0.5) / H * double(pi);
      if (theta <= maxThetaHorizon)
      {
        float shadow = float(std::min(std::max((maxThetaHorizon - theta) / (maxThetaHorizon - maxTheta), 0.), 1.));
        theta = std::min(theta, maxTheta);
        double phi = ((double(x) + 0.5) / W - 0.5) * (2.0 * double(pi));
        double cosGamma = std::cos(theta) * std::cos(sunTheta) + std::sin(theta) * std::sin(sunTheta) * std::cos(phi - sunPhi);
        double gamma = acosSafe(cosGamma);
        float radianceNorm = zero;
        for (int i = 0; i < cieSize; ++i)
          if (cieLambda[i] >= lambdaMin && cieLambda[i] <= lambdaMax)
          {
            double r = arhosekskymodel_radiance(model, theta, gamma, cieLambda[i]);
            radiance += float(r) * cieXyz(i);
            radianceNorm += cieY[i];
          }
        radiance = xyzToRgb(radiance / radianceNorm) * shadow;
      }
      radiance = max(radiance, Vec3f(0.0f));
      float* p = &sky[(size_t(y) * W + x) * 3];
      p[0] = radiance.x; p[1] = radiance.y; p[2] = radiance.z;
    }
  arhosekskymodelstate_free(model);

  FILE* f = fopen((out + "/sky.bin").c_str(), "wb");
  if (!f) { perror("sky.bin"); return 1; }
  fwrite(sky.data(), sizeof(float), sky.size(), f);
  fclose(f);
  f = fopen((out + "/sun.txt").c_str(), "w");
  fprintf(f, "%.9g %.9g %.9g %.9g %.9g %.9g %.9g\n", sunRadiance.x, sunRadiance.y, sunRadiance.z, sunRadius, sunDir.x, sunDir.y, sunDir.z);
  fclose(f);
  return 0;
}
