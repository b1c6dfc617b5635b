// Heterogeneous VQE batches (desv_batch_*): many (circuit, Hamiltonian,
// theta) instances of uneven qubit counts, evaluated together.
#include <hip/hip_runtime.h>

#include <chrono>
#include <cstring>
#include <memory>
#include <new>
#include <string>
#include <vector>

#include "../../include/desv.h"

namespace {

thread_local std::string g_berr;

struct Instance {
    int n = 0;
    std::vector<double> ops;        // [n_ops][4], angles patched from theta
    std::vector<int64_t> param_of_op;
    int64_t n_params = 0;
    std::vector<uint64_t> xmask, zmask;
    std::vector<double> coeff;
    desv_handle h = nullptr;
};

}  // namespace

struct desv_batch_s {
    int dtype = 0;
    int device = 0;
    int tile_bits = 0;
    int use_graph = 0;
    double last_us = 0;
    std::vector<Instance> inst;
};

namespace {
int bfail(int code, const std::string &m) {
    g_berr = m;
    return code;
}
}  // namespace

extern "C" {

int desv_batch_create(int dtype, int device, desv_batch *out) {
    if (!out) return DESV_EINVAL;
    if (dtype != DESV_C64 && dtype != DESV_C128) return DESV_EINVAL;
    desv_batch b = new (std::nothrow) desv_batch_s();
    if (!b) return DESV_ENOMEM;
    b->dtype = dtype;
    b->device = device;
    *out = b;
    return DESV_OK;
}

int desv_batch_destroy(desv_batch b) {
    if (!b) return DESV_OK;
    for (auto &i : b->inst)
        if (i.h) desv_destroy(i.h);
    delete b;
    return DESV_OK;
}

int desv_batch_add(desv_batch b, int n, const double *ops, int64_t n_ops, const int64_t *param_of_op,
                   int64_t n_params, const uint64_t *xmask, const uint64_t *zmask, const double *coeff,
                   int64_t n_terms, int64_t *index) {
    if (!b || n < 1 || n > 34 || n_ops < 0 || n_terms < 0 || n_params < 0) return DESV_EINVAL;
    if ((n_ops && (!ops || !param_of_op)) || (n_terms && (!xmask || !zmask || !coeff))) return DESV_EINVAL;
    Instance in;
    in.n = n;
    in.ops.assign(ops, ops + 4 * n_ops);
    in.param_of_op.assign(param_of_op, param_of_op + n_ops);
    for (int64_t i = 0; i < n_ops; ++i)
        if (in.param_of_op[i] >= n_params) return DESV_EINVAL;
    in.n_params = n_params;
    in.xmask.assign(xmask, xmask + n_terms);
    in.zmask.assign(zmask, zmask + n_terms);
    in.coeff.assign(coeff, coeff + n_terms);
    b->inst.push_back(std::move(in));
    if (index) *index = (int64_t)b->inst.size() - 1;
    return DESV_OK;
}

int desv_batch_set_params(desv_batch b, int64_t index, const double *theta) {
    if (!b || index < 0 || index >= (int64_t)b->inst.size()) return DESV_EINVAL;
    Instance &in = b->inst[index];
    if (in.n_params && !theta) return DESV_EINVAL;
    for (size_t i = 0; i < in.param_of_op.size(); ++i)
        if (in.param_of_op[i] >= 0) in.ops[4 * i + 3] = theta[in.param_of_op[i]];
    return DESV_OK;
}

int desv_batch_configure(desv_batch b, int tile_bits, int use_graph) {
    if (!b) return DESV_EINVAL;
    b->tile_bits = tile_bits;
    b->use_graph = use_graph;
    return DESV_OK;
}

int desv_batch_run(desv_batch b, double *energies, double *grads) {
    if (!b) return DESV_EINVAL;
    auto t0 = std::chrono::steady_clock::now();
    int64_t goff = 0;
    for (size_t i = 0; i < b->inst.size(); ++i) {
        Instance &in = b->inst[i];
        if (!in.h) {
            int rc = desv_create(in.n, b->dtype, b->device, &in.h);
            if (rc) return rc;
        } else {
            int rc = desv_reset(in.h);
            if (rc) return rc;
        }
        double e = 0;
        int rc = desv_adjoint(in.h, in.ops.data(), (int64_t)in.param_of_op.size(), in.param_of_op.data(),
                              in.n_params, in.xmask.data(), in.zmask.data(), in.coeff.data(),
                              (int64_t)in.xmask.size(), nullptr, &e, grads ? grads + goff : nullptr);
        if (rc) return rc;
        if (energies) energies[i] = e;
        goff += in.n_params;
    }
    b->last_us = std::chrono::duration<double, std::micro>(std::chrono::steady_clock::now() - t0).count();
    return DESV_OK;
}

int desv_batch_last_time_us(desv_batch b, double *us) {
    if (!b || !us) return DESV_EINVAL;
    *us = b->last_us;
    return DESV_OK;
}

}  // extern "C"
