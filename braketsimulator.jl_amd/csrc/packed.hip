// Compact op-list image: a bsim_op array as self-contained bytes (header,
// fixed records with value indices in place of pointers, complex128 values),
// and back without copying.  Host code only.
#include <stdint.h>
#include <string.h>

#include <vector>

#include "common.h"

namespace bsim {
namespace {

constexpr char kMagic[4] = { 'B', 'S', 'O', 'P' };
constexpr uint64_t kMaxKraus = 1u << 20;

// Complex values an op's matrix holds (false: counts out of range).
bool matrix_values(int type, int k, int n_kraus, uint64_t* out) {
    if (k < 1 || k > BSIM_MAX_TARGETS) return false;
    switch (type) {
        case BSIM_OP_GATE: *out = 1ull << (2 * k); return true;
        case BSIM_OP_DIAG: *out = 1ull << k; return true;
        case BSIM_OP_KRAUS_REF:
        case BSIM_OP_KRAUS_PHYS:
            if (n_kraus < 1 || (uint64_t)n_kraus > kMaxKraus) return false;
            *out = (uint64_t)n_kraus << (2 * k);
            return true;
        case BSIM_OP_PHASEFLIP: *out = 1; return true;
        case BSIM_OP_SUPEROP:
            if (k > 5) return false;
            *out = 1ull << (4 * k);
            return true;
        default: return false;
    }
}

// Derivative matrices (gates only) and the values each holds.
bool deriv_values(int type, int k, int n_params, int* count, uint64_t* each) {
    if (n_params < 0 || n_params > 3) return false;
    if (n_params > 0 && type != BSIM_OP_GATE && type != BSIM_OP_DIAG) return false;
    *count = n_params;
    *each = 1ull << (2 * k);
    return true;
}

bool counts_ok(const int32_t n_targets, const int32_t n_controls) {
    return n_targets >= 1 && n_targets <= BSIM_MAX_TARGETS && n_controls >= 0 && n_controls <= BSIM_MAX_CONTROLS;
}

}  // namespace
}  // namespace bsim

using namespace bsim;

extern "C" {

int bsim_ops_pack(const bsim_op* ops, int n_ops, void* buf, size_t cap, size_t* size) {
    if (n_ops < 0 || (n_ops > 0 && !ops) || !size) return fail(BSIM_EINVAL, "bad pack arguments");
    if (buf && ((uintptr_t)buf & 7)) return fail(BSIM_EINVAL, "pack buffer must be 8-byte aligned");
    uint64_t n_values = 0;
    for (int i = 0; i < n_ops; ++i) {
        const bsim_op& op = ops[i];
        uint64_t nv = 0, each = 0;
        int nd = 0;
        if (!counts_ok(op.n_targets, op.n_controls) || !matrix_values(op.type, op.n_targets, op.n_kraus, &nv) ||
            !deriv_values(op.type, op.n_targets, op.n_params, &nd, &each))
            return fail(BSIM_EINVAL, "pack: op " + std::to_string(i) + " has counts out of range");
        if (!op.matrix) return fail(BSIM_EINVAL, "pack: op " + std::to_string(i) + " has a null matrix");
        for (int j = 0; j < nd; ++j)
            if (!op.dmatrix[j]) return fail(BSIM_EINVAL, "pack: op " + std::to_string(i) + " has a null dmatrix");
        n_values += nv + (uint64_t)nd * each;
    }
    const size_t need = sizeof(bsim_packed_header) + sizeof(bsim_packed_op) * (size_t)n_ops + 16 * (size_t)n_values;
    *size = need;
    if (!buf) return BSIM_OK;
    if (cap < need) return fail(BSIM_EINVAL, "pack: buffer too small (" + std::to_string(need) + " bytes needed)");
    char* out = (char*)buf;
    bsim_packed_header hdr;
    memset(&hdr, 0, sizeof(hdr));
    memcpy(hdr.magic, kMagic, 4);
    hdr.version = BSIM_PACK_VERSION;
    hdr.n_ops = (uint32_t)n_ops;
    hdr.record_bytes = (uint32_t)sizeof(bsim_packed_op);
    hdr.n_values = n_values;
    memcpy(out, &hdr, sizeof(hdr));
    bsim_packed_op* rec = (bsim_packed_op*)(out + sizeof(hdr));
    double* val = (double*)(out + sizeof(hdr) + sizeof(bsim_packed_op) * (size_t)n_ops);
    uint64_t at = 0;
    for (int i = 0; i < n_ops; ++i) {
        const bsim_op& op = ops[i];
        bsim_packed_op r;
        memset(&r, 0, sizeof(r));
        r.type = op.type;
        r.n_targets = op.n_targets;
        r.n_controls = op.n_controls;
        r.n_kraus = op.n_kraus;
        memcpy(r.targets, op.targets, sizeof(r.targets));
        memcpy(r.controls, op.controls, sizeof(r.controls));
        memcpy(r.control_values, op.control_values, sizeof(r.control_values));
        r.n_params = op.n_params;
        memcpy(r.param_index, op.param_index, sizeof(r.param_index));
        uint64_t nv = 0, each = 0;
        int nd = 0;
        matrix_values(op.type, op.n_targets, op.n_kraus, &nv);
        deriv_values(op.type, op.n_targets, op.n_params, &nd, &each);
        r.matrix = (int64_t)at;
        if (op.type == BSIM_OP_PHASEFLIP) {  // the probability only: (p, 0)
            val[2 * at] = op.matrix[0];
            val[2 * at + 1] = 0.0;
        } else {
            memcpy(val + 2 * at, op.matrix, 16 * (size_t)nv);
        }
        at += nv;
        for (int j = 0; j < 3; ++j) r.dmatrix[j] = -1;
        for (int j = 0; j < nd; ++j) {
            r.dmatrix[j] = (int64_t)at;
            memcpy(val + 2 * at, op.dmatrix[j], 16 * (size_t)each);
            at += each;
        }
        memcpy(rec + i, &r, sizeof(r));
    }
    return BSIM_OK;
}

int bsim_ops_unpack(const void* buf, size_t size, bsim_op* ops, int cap, int* n_ops) {
    if (!buf || !n_ops) return fail(BSIM_EINVAL, "bad unpack arguments");
    if ((uintptr_t)buf & 7) return fail(BSIM_EINVAL, "unpack: image must be 8-byte aligned");
    if (size < sizeof(bsim_packed_header)) return fail(BSIM_EINVAL, "unpack: image shorter than its header");
    bsim_packed_header hdr;
    memcpy(&hdr, buf, sizeof(hdr));
    if (memcmp(hdr.magic, kMagic, 4) != 0) return fail(BSIM_EINVAL, "unpack: not an op-list image");
    if (hdr.version != BSIM_PACK_VERSION) return fail(BSIM_EINVAL, "unpack: unknown image version");
    if (hdr.record_bytes != sizeof(bsim_packed_op)) return fail(BSIM_EINVAL, "unpack: record size mismatch");
    const size_t body = size - sizeof(hdr);
    if (hdr.n_ops > body / sizeof(bsim_packed_op) || hdr.n_ops > (uint32_t)INT32_MAX)
        return fail(BSIM_EINVAL, "unpack: records run past the image");
    const size_t values_off = sizeof(hdr) + sizeof(bsim_packed_op) * (size_t)hdr.n_ops;
    if (hdr.n_values > (size - values_off) / 16) return fail(BSIM_EINVAL, "unpack: values run past the image");
    *n_ops = (int)hdr.n_ops;
    const char* in = (const char*)buf;
    const bsim_packed_op* rec = (const bsim_packed_op*)(in + sizeof(hdr));
    const double* val = (const double*)(in + values_off);
    auto inside = [&](int64_t at, uint64_t count) {
        return at >= 0 && (uint64_t)at <= hdr.n_values && count <= hdr.n_values - (uint64_t)at;
    };
    for (uint32_t i = 0; i < hdr.n_ops; ++i) {
        bsim_packed_op r;
        memcpy(&r, rec + i, sizeof(r));
        uint64_t nv = 0, each = 0;
        int nd = 0;
        if (!counts_ok(r.n_targets, r.n_controls) || !matrix_values(r.type, r.n_targets, r.n_kraus, &nv) ||
            !deriv_values(r.type, r.n_targets, r.n_params, &nd, &each))
            return fail(BSIM_EINVAL, "unpack: op " + std::to_string(i) + " has counts out of range");
        if (!inside(r.matrix, nv)) return fail(BSIM_EINVAL, "unpack: op " + std::to_string(i) + " matrix out of range");
        for (int j = 0; j < nd; ++j)
            if (!inside(r.dmatrix[j], each))
                return fail(BSIM_EINVAL, "unpack: op " + std::to_string(i) + " dmatrix out of range");
    }
    if (!ops) return BSIM_OK;
    if (cap < (int)hdr.n_ops) return fail(BSIM_EINVAL, "unpack: ops capacity too small");
    for (uint32_t i = 0; i < hdr.n_ops; ++i) {
        bsim_packed_op r;
        memcpy(&r, rec + i, sizeof(r));
        bsim_op& op = ops[i];
        memset(&op, 0, sizeof(op));
        op.type = r.type;
        op.n_targets = r.n_targets;
        op.n_controls = r.n_controls;
        op.n_kraus = r.n_kraus;
        memcpy(op.targets, r.targets, sizeof(op.targets));
        memcpy(op.controls, r.controls, sizeof(op.controls));
        memcpy(op.control_values, r.control_values, sizeof(op.control_values));
        op.matrix = val + 2 * r.matrix;
        op.n_params = r.n_params;
        memcpy(op.param_index, r.param_index, sizeof(op.param_index));
        for (int j = 0; j < r.n_params; ++j) op.dmatrix[j] = val + 2 * r.dmatrix[j];
    }
    return BSIM_OK;
}

int bsim_evolve_packed(bsim_state* h, const void* buf, size_t size, int mode) {
    if (int rc = check_state(h)) return rc;
    int n = 0;
    if (int rc = bsim_ops_unpack(buf, size, nullptr, 0, &n)) return rc;
    std::vector<bsim_op> ops((size_t)std::max(n, 1));
    if (int rc = bsim_ops_unpack(buf, size, ops.data(), n, &n)) return rc;
    return bsim_evolve(h, ops.data(), n, mode);
}

}  // extern "C"
