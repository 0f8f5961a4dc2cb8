// hipGraph capture of sort launch sequences.
//
// Every sort entry point is stream-ordered and issues only kernel launches (no host
// synchronisation, no memset node), so a caller can record a whole call -- or a batch
// of calls -- once and replay it with one graph launch.  These helpers wrap HIP stream
// capture so C and Python callers do not need the HIP graph API themselves.
#include <hip/hip_runtime.h>

#include "../../include/cccl_amd/radix_sort_amd.h"

extern "C" {

CCCL_C_API int cccl_amd_capture_begin(CUstream stream)
{
  if (stream == nullptr)
  {
    return hipErrorInvalidValue; // the legacy default stream cannot be captured
  }
  return int(hipStreamBeginCapture(reinterpret_cast<hipStream_t>(stream), hipStreamCaptureModeThreadLocal));
}

CCCL_C_API int cccl_amd_capture_end(CUstream stream, void** graph_exec)
{
  if (stream == nullptr || graph_exec == nullptr)
  {
    return hipErrorInvalidValue;
  }
  *graph_exec     = nullptr;
  hipGraph_t g    = nullptr;
  hipError_t e    = hipStreamEndCapture(reinterpret_cast<hipStream_t>(stream), &g);
  if (e != hipSuccess)
  {
    return int(e);
  }
  hipGraphExec_t x = nullptr;
  e                = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
  (void) hipGraphDestroy(g);
  if (e != hipSuccess)
  {
    return int(e);
  }
  *graph_exec = x;
  return hipSuccess;
}

CCCL_C_API int cccl_amd_graph_launch(void* graph_exec, CUstream stream)
{
  if (graph_exec == nullptr)
  {
    return hipErrorInvalidValue;
  }
  return int(hipGraphLaunch(static_cast<hipGraphExec_t>(graph_exec), reinterpret_cast<hipStream_t>(stream)));
}

CCCL_C_API int cccl_amd_graph_destroy(void* graph_exec)
{
  if (graph_exec == nullptr)
  {
    return hipErrorInvalidValue;
  }
  return int(hipGraphExecDestroy(static_cast<hipGraphExec_t>(graph_exec)));
}

} // extern "C"
