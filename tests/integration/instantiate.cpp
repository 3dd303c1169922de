// Compiles only if DeviceHip and HipDenoiser implement every pure virtual of the reference's
// Device (render/device.h:56-114) and Denoiser (image/denoiser.h:10-41): both are instantiated.
#include <memory>
#include <vector>

#include "device_hip.h"
#include "hip_denoiser.h"

std::shared_ptr<prt::Device> makeHipDevice() { return std::make_shared<prt::DeviceHip>(0); }

std::shared_ptr<prt::Denoiser> makeHipDenoiser(const std::vector<float>& weights)
{
  return std::make_shared<prt::HipDenoiser>(0, prt::HipDenoiser::WeightSets{weights, {}, {}});
}
