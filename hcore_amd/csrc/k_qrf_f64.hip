// k_qrf_f64.hip — double instantiation of launch_qr_fused (kernels in qr_reg.hpp).
#include "qr_reg.hpp"

namespace hcore {
template bool launch_qr_fused<double>(const QrProb<double>*, int, int, hipStream_t, const int*, int);
}  // namespace hcore
