// unipi::axhel drop-in API (include/axhel/*.hpp) over the gfx950 C ABI.
//
// Same signatures and semantics as the reference's public functions
// (axhel/eltwise/eltwise-*-mod.cpp dispatchers, axhel/ntt/ntt.cpp:124-216):
// every call returns with the result in place, unsupported mod factors leave
// the output untouched without error, and results may alias operands.  Host
// and device pointers are both accepted.  Device failures, which the
// reference cannot have, surface as std::runtime_error.
#include <stdexcept>
#include <string>

#include "axhel/axhel.hpp"
#include "axhel_hip.h"

namespace unipi {
namespace axhel {
namespace {

void check(axhel_hip_status s, const char *what) {
    if (s == AXHEL_HIP_OK || s == AXHEL_HIP_UNSUPPORTED_FACTOR) return;
    std::string msg = std::string("axhel: ") + what + ": " + axhel_hip_status_string(s);
    const char *detail = axhel_hip_last_error();
    if (detail && *detail) msg += std::string(" (") + detail + ")";
    if (s == AXHEL_HIP_INVALID_ARGUMENT) throw std::invalid_argument(msg);
    throw std::runtime_error(msg);
}

// Device-pointer calls are asynchronous on the default stream; the drop-in
// contract is synchronous.
void done(axhel_hip_status s, const char *what) {
    check(s, what);
    if (s == AXHEL_HIP_OK) check(axhel_hip_stream_synchronize(nullptr), what);
}

}  // namespace

void EltwiseAddMod(uint64_t *result, const uint64_t *operand1, const uint64_t *operand2, uint64_t n,
                   uint64_t modulus) {
    done(axhel_hip_add_mod_vv(result, operand1, operand2, n, modulus, nullptr), "EltwiseAddMod");
}

void EltwiseAddMod(uint64_t *result, const uint64_t *operand1, uint64_t operand2, uint64_t n, uint64_t modulus) {
    done(axhel_hip_add_mod_vs(result, operand1, operand2, n, modulus, nullptr), "EltwiseAddMod");
}

void EltwiseSubMod(uint64_t *result, const uint64_t *operand1, const uint64_t *operand2, uint64_t n,
                   uint64_t modulus) {
    done(axhel_hip_sub_mod_vv(result, operand1, operand2, n, modulus, nullptr), "EltwiseSubMod");
}

void EltwiseSubMod(uint64_t *result, const uint64_t *operand1, uint64_t operand2, uint64_t n, uint64_t modulus) {
    done(axhel_hip_sub_mod_vs(result, operand1, operand2, n, modulus, nullptr), "EltwiseSubMod");
}

void EltwiseMulMod(uint64_t *result, const uint64_t *operand1, const uint64_t *operand2, uint64_t n, uint64_t modulus,
                   uint64_t input_mod_factor) {
    done(axhel_hip_mul_mod(result, operand1, operand2, n, modulus, input_mod_factor, nullptr), "EltwiseMulMod");
}

void EltwiseFMAMod(uint64_t *res, const uint64_t *op1, uint64_t op2, const uint64_t *op3, uint64_t n, uint64_t mod,
                   uint64_t mod_factor) {
    done(axhel_hip_fma_mod(res, op1, op2, op3, n, mod, mod_factor, nullptr), "EltwiseFMAMod");
}

void EltwiseReduceMod(uint64_t *res, const uint64_t *op, uint64_t n, uint64_t mod, uint64_t in_mod_factor,
                      uint64_t out_mod_factor) {
    done(axhel_hip_reduce_mod(res, op, n, mod, in_mod_factor, out_mod_factor, nullptr), "EltwiseReduceMod");
}

// ntt.cpp:25-40: the shared implementation; here it owns a device plan.
class NTT::Impl {
public:
    axhel_hip_plan *plan = nullptr;
    uint64_t degree = 0;
    uint64_t modulus = 0;
    ~Impl() { axhel_hip_ntt_plan_destroy(plan); }
};

NTT::NTT(uint64_t degree, uint64_t modulus, uint64_t root_of_unity) : impl_(std::make_shared<Impl>()) {
    impl_->degree = degree;
    impl_->modulus = modulus;
    check(axhel_hip_ntt_plan_create(degree, modulus, root_of_unity, &impl_->plan), "NTT");
}

NTT::NTT(uint64_t degree, uint64_t modulus, const NTTMultiplyOperand *root_powers,
         const NTTMultiplyOperand *inv_root_powers, NTTMultiplyOperand inv_degree_modulo)
    : impl_(std::make_shared<Impl>()) {
    static_assert(sizeof(NTTMultiplyOperand) == 2 * sizeof(uint64_t), "AoS table layout");
    impl_->degree = degree;
    impl_->modulus = modulus;
    check(axhel_hip_ntt_plan_create_external(degree, modulus, reinterpret_cast<const uint64_t *>(root_powers),
                                             reinterpret_cast<const uint64_t *>(inv_root_powers),
                                             inv_degree_modulo.operand, inv_degree_modulo.quotient, &impl_->plan),
          "NTT");
}

NTT::~NTT() = default;
NTT::NTT(const NTT &) noexcept = default;
NTT &NTT::operator=(const NTT &) noexcept = default;
NTT::NTT(NTT &&) noexcept = default;
NTT &NTT::operator=(NTT &&) noexcept = default;

void NTT::ComputeForward(uint64_t *result, const uint64_t *operand, uint64_t input_mod_factor,
                         uint64_t output_mod_factor) const {
    done(axhel_hip_ntt_forward(impl_->plan, result, operand, input_mod_factor, output_mod_factor, nullptr),
         "NTT::ComputeForward");
}

void NTT::ComputeInverse(uint64_t *result, const uint64_t *operand, uint64_t input_mod_factor,
                         uint64_t output_mod_factor) const {
    done(axhel_hip_ntt_inverse(impl_->plan, result, operand, input_mod_factor, output_mod_factor, nullptr),
         "NTT::ComputeInverse");
}

uint64_t NTT::Degree() const noexcept { return impl_ ? impl_->degree : 0; }

uint64_t NTT::Modulus() const noexcept { return impl_ ? impl_->modulus : 0; }

}  // namespace axhel
}  // namespace unipi
