// x2_host.h — host helpers shared by the library's host files, and the kernel launchers they call.
#pragma once
#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "x2_layout.h"

namespace x2 {
void sobolTable(uint32_t out[17 * 32]);               // the Sobol direction numbers (x2_lane.hip)
// The lights' device buffer: the x2_light records, then 12 floats a light (lightCorners). Its bytes, and its contents.
size_t lightBufferBytes(uint32_t count);
std::vector<char> lightBuffer(const x2_light* lights, uint32_t count);
int reportFailure(int code, const std::string& what);  // sets x2dev_last_error's text, returns code
hipError_t uploadSobol(const uint32_t* table);
hipError_t batchResident(int waves, uint32_t threads, int& blocks);
hipError_t launchBatch(const BatchParams& b, int waves, uint32_t threads, int blocks, hipStream_t stream);
hipError_t launchFill(const FillRegion* regions, uint32_t n, hipStream_t stream);
hipError_t launchStamp(unsigned long long* at, hipStream_t stream);
}  // namespace x2
