// HipDenoiser: the learned denoiser (BF16-MFMA U-Net on an MI355X) behind ProtoRay's Denoiser
// interface (image/denoiser.h:10-41), beside OIDNDenoiser (image/oidn_denoiser.cpp:10-130).  The
// maintainer creates it where the reference creates its denoiser and gives it the trained weights
// of each input set, as OIDN keeps weights per input set: colour + albedo + normal
// (oidn-protoray_amd/weights/unet_sunsky.npy), colour + albedo (unet_sunsky_color_albedo.npy) and
// colour only (unet_sunsky_color.npy), float32 values.  init picks the set the Params name (albedo
// and normal may be null, a normal needs an albedo); the missing images enter as zero channels.
#pragma once

#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "image/denoiser.h"
#include "prt_hip.h"

namespace prt {

class HipDenoiser : public Denoiser
{
public:
  struct WeightSets
  {
    std::vector<float> full;         // colour, albedo, normal
    std::vector<float> colorAlbedo;  // colour, albedo
    std::vector<float> color;        // colour only
  };

  HipDenoiser(int deviceOrdinal, WeightSets weightSets) : sets(std::move(weightSets))
  {
    check(prt_denoiser_create(deviceOrdinal, &dn));
    int64_t n = 0;
    check(prt_denoiser_param_count(dn, &n));
    for (const std::vector<float>* w : {&sets.full, &sets.colorAlbedo, &sets.color})
      if (!w->empty() && int64_t(w->size()) != n)
      {
        prt_denoiser_destroy(dn);
        throw std::invalid_argument("denoiser weights: expected " + std::to_string(n) + " values per set");
      }
  }
  ~HipDenoiser() override { prt_denoiser_destroy(dn); }
  HipDenoiser(const HipDenoiser&) = delete;
  HipDenoiser& operator =(const HipDenoiser&) = delete;

  // Denoiser::init: the images of Params (H x W x 3 floats each, output written by execute)
  void init(const Params& params) override
  {
    if (params.srgb) throw std::invalid_argument("HipDenoiser filters linear images only (srgb = false)");
    const std::vector<float>& w = params.normal ? sets.full : params.albedo ? sets.colorAlbedo : sets.color;
    if (w.empty()) throw std::invalid_argument("no denoiser weights for this combination of input images");
    if (&w != current)
    {
      check(prt_denoiser_set_weights(dn, w.data(), int64_t(w.size())));
      current = &w;
    }
    prt_denoiser_params p{};
    p.color = params.color;
    p.albedo = params.albedo;
    p.normal = params.normal;
    p.output = params.output;
    p.width = params.width;
    p.height = params.height;
    p.hdr = params.hdr ? 1 : 0;
    p.hdrScale = params.hdrScale;  // NaN: 1, OIDN's default
    check(prt_denoiser_init(dn, &p));
  }

  // Denoiser::execute: returns the denoising time in seconds
  double execute(bool benchmark = false) override
  {
    (void)benchmark;
    double seconds = 0.0;
    check(prt_denoiser_execute(dn, &seconds));
    return seconds;
  }

  std::string getVersion() override
  {
    char buf[128];
    check(prt_denoiser_get_version(dn, buf, int32_t(sizeof(buf))));
    return buf;
  }

  prt_denoiser* getHandle() const { return dn; }

private:
  static void check(int rc)
  {
    if (rc == PRT_OK) return;
    if (rc == PRT_ERR_INVALID || rc == PRT_ERR_STATE) throw std::invalid_argument(prt_last_error());
    throw std::runtime_error(prt_last_error());
  }

  prt_denoiser* dn = nullptr;
  WeightSets sets;
  const std::vector<float>* current = nullptr;  // the set on the device
};

} // namespace prt
