// Reference include path <states/SingleArmStates.h>: forwards to the CUDAMPLib API over the MI355X C ABI.
#pragma once
#include "../CUDAMPLib/mp_cudampl.h"
