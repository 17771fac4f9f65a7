// CUDAMPLib utilities (cudampl/CUDAMPLib/include/util.h), over the MI355X C ABI (include/mp_abi.h).
#pragma once
#include <vector>

namespace CUDAMPLib {

std::vector<std::vector<float>> interpolateVectors(const std::vector<float>& v1, const std::vector<float>& v2,
                                                   float resolution);
// vectors of vec without those equal (componentwise within tol) to an earlier kept one
// (util.cpp:250-267)
std::vector<std::vector<float>> removeRedundantVectors(const std::vector<std::vector<float>>& vec, float tol);

}  // namespace CUDAMPLib
