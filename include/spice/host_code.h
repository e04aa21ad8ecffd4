// Ends the host+device region that spice/snn.h opens for user model structs.
// Include it after the models in a translation unit that goes on to define
// host-only functions with external linkage (other than main).
#pragma once
#pragma clang force_cuda_host_device end
