// CUDAMPLib planner API (cudampl/CUDAMPLib/include/{base,spaces,states,constraints,tasks,
// termination,planners}) re-implemented over the MI355X C ABI.  Class names, constructor
// arguments and the methods the reference's callers use keep their reference signatures; the
// per-constraint CUDA kernels, NVRTC and the host-side graph code are replaced by the fused HIP
// kernels behind include/mp_abi.h.
//
// The headers sit where the reference installs them (include/CUDAMPLib/<area>/<Class>.h,
// cudampl/CUDAMPLib/CMakeLists.txt:57), so callers' <CUDAMPLib/spaces/SingleArmSpace.h> includes
// resolve with this directory's parent on the include path.  This header includes all of them.
#pragma once

#include "base/Constraint.h"
#include "base/Planner.h"
#include "base/Space.h"
#include "base/States.h"
#include "base/Task.h"
#include "base/Termination.h"
#include "constraints/BoundaryConstraint.h"
#include "constraints/EnvConstraintCuboid.h"
#include "constraints/EnvConstraintCylinder.h"
#include "constraints/EnvConstraintSphere.h"
#include "constraints/SelfCollisionConstraint.h"
#include "planners/RRG.h"
#include "planners/cRRTC.h"
#include "spaces/SingleArmSpace.h"
#include "states/SingleArmStates.h"
#include "tasks/SingleArmTask.h"
#include "termination/StepTermination.h"
#include "termination/TimeoutTermination.h"
#include "util.h"
