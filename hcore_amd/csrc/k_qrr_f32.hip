// k_qrr_f32.hip — float instantiation of launch_qr_reg (kernels in qr_reg.hpp).
#include "qr_reg.hpp"

namespace hcore {
template bool launch_qr_reg<float>(const QrProb<float>*, int, int, int, hipStream_t, const int*, int);
}  // namespace hcore
