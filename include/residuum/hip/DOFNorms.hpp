// residuum/hip/DOFNorms.hpp -- per-DOF residual norms of device vectors.
//
// The reference's loggers and NewtonRunner (solver/nonlinear/Newton.hpp:49)
// take a residual's per-field 2-norms from its raw data() pointer
// (fem::dof::computePerDOFNorms, fem/dof/DOFResidualNorms.hpp:14-53), which
// for a Vector<Real, Hip> is device memory.  This overload for double
// residuals, preferred over that template wherever it is declared first
// (include this header before the reference's solver and logging headers),
// copies device data to the host and then forms the norms as the
// reference defines them: per component the sum of r_i^2 in index order
// (interleaved: component i % numDOFs; block: the component's contiguous
// range), then the square root.  Host data is used in place, so CPU vectors
// get exactly the reference's values.
#pragma once

#include <cmath>
#include <vector>

#include "core/Types.hpp"
#include "fem/dof/DOFOrdering.hpp"
#include "residuum/hip/Backend.hpp"

namespace residuum {
namespace fem {
namespace dof {

inline std::vector<double> computePerDOFNorms(const double* r, Index totalSize, Index numDOFs, Index freeDOFsPerField,
                                              DOFOrdering ordering) {
    int32_t dev = 0;
    hip::check(rsd_pointer_is_device(r, &dev));
    std::vector<double> host;
    if (dev && totalSize) {
        host.resize(totalSize);
        linalg::types::backend::Hip::copy<double>(host.data(), r, totalSize, linalg::types::backend::CopyKind::DeviceToHost);
        r = host.data();
    }
    const Index nf = numDOFs > 0 ? numDOFs : 1;
    std::vector<double> sq(nf, 0.0);
    if (numDOFs <= 1 || freeDOFsPerField == 0) {
        for (Index i = 0; i < totalSize; ++i) sq[0] += r[i] * r[i];
        return {std::sqrt(sq[0])};
    }
    if (ordering == DOFOrdering::Interleaved) {
        for (Index i = 0; i < totalSize; ++i) sq[i % numDOFs] += r[i] * r[i];
    } else {
        for (Index c = 0; c < numDOFs; ++c)
            for (Index i = c * freeDOFsPerField; i < (c + 1) * freeDOFsPerField && i < totalSize; ++i) sq[c] += r[i] * r[i];
    }
    for (double& v : sq) v = std::sqrt(v);
    return sq;
}

}  // namespace dof
}  // namespace fem
}  // namespace residuum
