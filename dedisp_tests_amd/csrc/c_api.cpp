// C API of libdedisp.so: reference src/clib/dedisp.cpp.
//
// Same entry points, arguments and return conventions. Differences, all
// within the documented contract of src/clib/dedisp.h:
//  - the specific dedisp_error of a failure is returned (the reference maps
//    every C++ exception to DEDISP_UNKNOWN_ERROR, dedisp.cpp:154-157, 520-523);
//  - DEDISP_DEVICE_POINTERS and DEDISP_ASYNC are honoured;
//  - DEDISP_DEDISP (the legacy implementation) runs the TDD kernels, which
//    compute the same per-sample arithmetic.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <memory>
#include <new>

#include "internal.hpp"

namespace {

int g_device_idx = 0;                                  // dedisp.cpp:31
dedisp_implementation g_implementation = DEDISP_DEDISP;  // dedisp.cpp:34

// Entry check of every reference C function: a pending GPU error from an
// earlier call is reported as DEDISP_PRIOR_GPU_ERROR (dedisp.cpp:117, 167, ...).
bool prior_gpu_error()
{
    hipError_t e = hipGetLastError();
    return e != hipSuccess && e != hipErrorNoDevice;
}

template <typename F>
dedisp_error guarded(F&& f)
{
    try {
        f();
    } catch (const dedisp::Error& e) {
        return e.code();
    } catch (const std::bad_alloc&) {
        return DEDISP_MEM_ALLOC_FAILED;
    } catch (...) {
        return DEDISP_UNKNOWN_ERROR;
    }
    return DEDISP_NO_ERROR;
}

dedisp::TDDPlan* as_tdd(dedisp_plan plan) { return dynamic_cast<dedisp::TDDPlan*>(plan->ptr.get()); }
dedisp::FDDGPUPlan* as_fdd(dedisp_plan plan) { return dynamic_cast<dedisp::FDDGPUPlan*>(plan->ptr.get()); }
dedisp::GPUPlan* as_gpu(dedisp_plan plan) { return dynamic_cast<dedisp::GPUPlan*>(plan->ptr.get()); }

}  // namespace

extern "C" {

dedisp_error dedisp_create_plan(dedisp_plan* plan, dedisp_size nchans, dedisp_float dt,
                                dedisp_float f0, dedisp_float df)
{
    if (!plan) return DEDISP_INVALID_POINTER;
    *plan = nullptr;
    if (prior_gpu_error()) return DEDISP_PRIOR_GPU_ERROR;

    // USE_CPU selects the reference's CPU FDD plan (dedisp.cpp:129-146). This
    // library runs every plan on the GPU; the CPU FDD algorithm is not part of
    // it, so the request is refused rather than silently served by the GPU.
    const char* use_cpu_str = std::getenv("USE_CPU");
    const bool use_cpu = use_cpu_str ? std::atoi(use_cpu_str) != 0 : false;
    if (use_cpu && g_implementation == DEDISP_FDD) {
        std::fprintf(stderr, "dedisp: USE_CPU=1 requests the CPU FDD plan, which this "
                             "GPU library does not provide\n");
        return DEDISP_UNKNOWN_ERROR;
    }

    dedisp_plan p = new (std::nothrow) dedisp_plan_struct();
    if (!p) return DEDISP_MEM_ALLOC_FAILED;
    dedisp_error err = guarded([&] {
        switch (g_implementation) {
        case DEDISP_DEDISP:
            p->ptr.reset(new dedisp::DedispPlan(nchans, dt, f0, df, g_device_idx));
            break;
        case DEDISP_TDD:
            p->ptr.reset(new dedisp::TDDPlan(nchans, dt, f0, df, g_device_idx));
            break;
        case DEDISP_FDD:
            p->ptr.reset(new dedisp::FDDGPUPlan(nchans, dt, f0, df, g_device_idx));
            break;
        default:
            DD_THROW(DEDISP_UNKNOWN_ERROR);
        }
        p->impl = g_implementation;
    });
    if (err != DEDISP_NO_ERROR) {
        delete p;
        return err;
    }
    *plan = p;
    return DEDISP_NO_ERROR;
}

void dedisp_destroy_plan(dedisp_plan plan)
{
    if (plan) delete plan;
}

dedisp_error dedisp_set_device(int device_idx)
{
    g_device_idx = device_idx;
    return DEDISP_NO_ERROR;
}

dedisp_error dedisp_select_implementation(dedisp_implementation imp)
{
    if (imp != DEDISP_DEDISP && imp != DEDISP_TDD && imp != DEDISP_FDD) return DEDISP_UNKNOWN_ERROR;
    g_implementation = imp;
    return DEDISP_NO_ERROR;
}

dedisp_error dedisp_set_gulp_size(dedisp_plan plan, dedisp_size gulp_size)
{
    if (!plan) return DEDISP_INVALID_PLAN;
    if (prior_gpu_error()) return DEDISP_PRIOR_GPU_ERROR;
    dedisp::TDDPlan* t = as_tdd(plan);
    if (!t) return DEDISP_INVALID_PLAN;  // FDD plans have no gulp (dedisp.cpp:182-183)
    return guarded([&] { t->set_gulp_size(gulp_size); });
}

dedisp_size dedisp_get_gulp_size(dedisp_plan plan)
{
    if (!plan) return 0;
    dedisp::TDDPlan* t = as_tdd(plan);
    if (!t) return 0;
    return t->get_gulp_size();
}

dedisp_error dedisp_set_dm_list(dedisp_plan plan, const dedisp_float* dm_list, dedisp_size count)
{
    if (!plan) return DEDISP_INVALID_PLAN;
    if (prior_gpu_error()) return DEDISP_PRIOR_GPU_ERROR;
    return guarded([&] { plan->ptr->set_dm_list(dm_list, count); });
}

dedisp_error dedisp_generate_dm_list(dedisp_plan plan, dedisp_float dm_start, dedisp_float dm_end,
                                     dedisp_float ti, dedisp_float tol)
{
    if (!plan) return DEDISP_INVALID_PLAN;
    if (prior_gpu_error()) return DEDISP_PRIOR_GPU_ERROR;
    return guarded([&] { plan->ptr->generate_dm_list(dm_start, dm_end, ti, tol); });
}

dedisp_float* dedisp_generate_dm_list_guru(dedisp_float dm_start, dedisp_float dm_end, double dt,
                                           double ti, double f0, double df, dedisp_size nchans,
                                           double tol, dedisp_size* dm_count)
{
    if (!dm_count) return nullptr;
    *dm_count = 0;
    try {
        std::vector<dedisp_float> table;
        dedisp::Plan::generate_dm_list(table, dm_start, dm_end, dt, ti, f0, df, nchans, tol);
        dedisp_float* outp = static_cast<dedisp_float*>(std::malloc(table.size() * sizeof(dedisp_float)));
        if (!outp) return nullptr;
        std::copy(table.begin(), table.end(), outp);
        *dm_count = table.size();
        return outp;
    } catch (...) {
        return nullptr;
    }
}

void dedisp_free_dm_list_guru(dedisp_float* dm_list) { std::free(dm_list); }

dedisp_error dedisp_set_killmask(dedisp_plan plan, const dedisp_bool* killmask)
{
    if (!plan) return DEDISP_INVALID_PLAN;
    if (prior_gpu_error()) return DEDISP_PRIOR_GPU_ERROR;
    return guarded([&] { plan->ptr->set_killmask(killmask); });
}

// ---- getters (dedisp.cpp:321-370): 0 on error ----

dedisp_size dedisp_get_max_delay(const dedisp_plan plan)
{
    if (!plan) return 0;
    if (plan->ptr->get_dm_count() == 0) return 0;
    return (dedisp_size)plan->ptr->get_max_delay();
}

dedisp_size dedisp_get_dm_delay(const dedisp_plan plan, int dm_trial)
{
    if (!plan) return 0;
    if (plan->ptr->get_dm_count() == 0) return 0;
    if (dm_trial < 0 || (dedisp_size)dm_trial >= plan->ptr->get_dm_count()) return 0;
    // dedisp.cpp:333: float product, + 0.5 in double, truncated
    return (dedisp_size)(plan->ptr->get_dm_list()[dm_trial] *
                             plan->ptr->get_delay_table()[plan->ptr->get_channel_count() - 1] +
                         0.5);
}

dedisp_size dedisp_get_channel_count(const dedisp_plan plan)
{
    return plan ? plan->ptr->get_channel_count() : 0;
}

dedisp_size dedisp_get_dm_count(const dedisp_plan plan)
{
    return plan ? plan->ptr->get_dm_count() : 0;
}

const dedisp_float* dedisp_get_dm_list(const dedisp_plan plan)
{
    if (!plan) return nullptr;
    if (plan->ptr->get_dm_count() == 0) return nullptr;
    return plan->ptr->get_dm_list();
}

const dedisp_bool* dedisp_get_killmask(const dedisp_plan plan)
{
    return plan ? plan->ptr->get_killmask() : nullptr;
}

dedisp_float dedisp_get_dt(const dedisp_plan plan) { return plan ? plan->ptr->get_dt() : 0.f; }
dedisp_float dedisp_get_f0(const dedisp_plan plan) { return plan ? plan->ptr->get_f0() : 0.f; }
dedisp_float dedisp_get_df(const dedisp_plan plan) { return plan ? plan->ptr->get_df() : 0.f; }

// Strings of src/common/dedisp_error.cpp:3-42.
const char* dedisp_get_error_string(dedisp_error error)
{
    switch (error) {
    case DEDISP_NO_ERROR: return "No error";
    case DEDISP_MEM_ALLOC_FAILED: return "Memory allocation failed";
    case DEDISP_MEM_COPY_FAILED: return "Memory copy failed";
    case DEDISP_INVALID_DEVICE_INDEX: return "Invalid device index";
    case DEDISP_DEVICE_ALREADY_SET: return "Device is already set and cannot be changed";
    case DEDISP_NCHANS_EXCEEDS_LIMIT: return "No. channels exceeds internal limit";
    case DEDISP_INVALID_PLAN: return "Invalid plan";
    case DEDISP_INVALID_POINTER: return "Invalid pointer";
    case DEDISP_INVALID_STRIDE: return "Invalid stride";
    case DEDISP_NO_DM_LIST_SET: return "No DM list has been set";
    case DEDISP_TOO_FEW_NSAMPS: return "No. samples < maximum delay";
    case DEDISP_INVALID_FLAG_COMBINATION: return "Invalid flag combination";
    case DEDISP_UNSUPPORTED_IN_NBITS: return "Unsupported in_nbits value";
    case DEDISP_UNSUPPORTED_OUT_NBITS: return "Unsupported out_nbits value";
    case DEDISP_PRIOR_GPU_ERROR: return "Prior GPU error.";
    case DEDISP_INTERNAL_GPU_ERROR: return "Internal GPU error. Please contact the author(s).";
    case DEDISP_UNKNOWN_ERROR: return "Unknown error. Please contact the author(s).";
    default: return "Invalid error code";
    }
}

// ---- execution ----

dedisp_error dedisp_execute_guru(const dedisp_plan plan, dedisp_size nsamps, const dedisp_byte* in,
                                 dedisp_size in_nbits, dedisp_size in_stride, dedisp_byte* out,
                                 dedisp_size out_nbits, dedisp_size out_stride,
                                 dedisp_size first_dm_idx, dedisp_size dm_count, unsigned flags)
{
    if (!plan) return DEDISP_INVALID_PLAN;
    if (prior_gpu_error()) return DEDISP_PRIOR_GPU_ERROR;
    dedisp::TDDPlan* t = as_tdd(plan);
    if (!t) return DEDISP_INVALID_PLAN;  // dedisp.cpp:428-429
    return guarded([&] {
        t->execute_guru(nsamps, in, in_nbits, in_stride, out, out_nbits, out_stride, first_dm_idx,
                        dm_count, flags);
    });
}

dedisp_error dedisp_execute_adv(const dedisp_plan plan, dedisp_size nsamps, const dedisp_byte* in,
                                dedisp_size in_nbits, dedisp_size in_stride, dedisp_byte* out,
                                dedisp_size out_nbits, dedisp_size out_stride, unsigned flags)
{
    if (!plan) return DEDISP_INVALID_PLAN;
    if (prior_gpu_error()) return DEDISP_PRIOR_GPU_ERROR;
    dedisp::TDDPlan* t = as_tdd(plan);
    if (!t) return DEDISP_INVALID_PLAN;  // dedisp.cpp:488-489
    return guarded([&] {
        t->execute_adv(nsamps, in, in_nbits, in_stride, out, out_nbits, out_stride, flags);
    });
}

dedisp_error dedisp_execute(const dedisp_plan plan, dedisp_size nsamps, const dedisp_byte* in,
                            dedisp_size in_nbits, dedisp_byte* out, dedisp_size out_nbits,
                            unsigned flags)
{
    if (!plan) return DEDISP_INVALID_PLAN;
    if (prior_gpu_error()) return DEDISP_PRIOR_GPU_ERROR;
    return guarded([&] { plan->ptr->execute(nsamps, in, in_nbits, out, out_nbits, flags); });
}

dedisp_error dedisp_amd_execute_batch(dedisp_plan plan, dedisp_size nbeams, const dedisp_size* nsamps,
                                      const dedisp_byte* const* in, dedisp_size in_nbits,
                                      dedisp_byte* const* out, dedisp_size out_nbits, unsigned flags)
{
    if (!plan) return DEDISP_INVALID_PLAN;
    if (prior_gpu_error()) return DEDISP_PRIOR_GPU_ERROR;
    if (nbeams && (!nsamps || !in || !out)) return DEDISP_INVALID_POINTER;
    if (dedisp::TDDPlan* t = as_tdd(plan))
        return guarded([&] { t->execute_batch(nbeams, nsamps, in, in_nbits, out, out_nbits, flags); });
    if (dedisp::FDDGPUPlan* f = as_fdd(plan))
        return guarded([&] { f->execute_batch(nbeams, nsamps, in, in_nbits, out, out_nbits, flags); });
    // other plans: one beam after another
    return guarded([&] {
        for (dedisp_size b = 0; b < nbeams; ++b)
            plan->ptr->execute(nsamps[b], in[b], in_nbits, out[b], out_nbits, flags);
    });
}

dedisp_error dedisp_amd_autotune(dedisp_plan plan, dedisp_size nsamps, dedisp_size in_nbits,
                                 dedisp_size out_nbits, dedisp_amd_tuning* result)
{
    if (!plan) return DEDISP_INVALID_PLAN;
    if (prior_gpu_error()) return DEDISP_PRIOR_GPU_ERROR;
    if (dedisp::TDDPlan* t = as_tdd(plan))
        return guarded([&] { t->autotune(nsamps, in_nbits, out_nbits, result); });
    if (dedisp::FDDGPUPlan* f = as_fdd(plan))
        return guarded([&] { f->autotune(nsamps, in_nbits, out_nbits, result); });
    return DEDISP_INVALID_PLAN;
}

dedisp_error dedisp_sync(void)
{
    if (hipDeviceSynchronize() != hipSuccess) {
        (void)hipGetLastError();
        return DEDISP_PRIOR_GPU_ERROR;
    }
    return DEDISP_NO_ERROR;
}

// ---- adaptive dt: declared by the reference, never implemented ----

dedisp_error dedisp_enable_adaptive_dt(dedisp_plan plan, dedisp_float, dedisp_float)
{
    if (!plan) return DEDISP_INVALID_PLAN;
    return DEDISP_UNKNOWN_ERROR;
}

dedisp_error dedisp_disable_adaptive_dt(dedisp_plan plan)
{
    return plan ? DEDISP_NO_ERROR : DEDISP_INVALID_PLAN;
}

dedisp_bool dedisp_using_adaptive_dt(const dedisp_plan) { return 0; }

const dedisp_size* dedisp_get_dt_factors(const dedisp_plan) { return nullptr; }

// ---- extensions (dedisp_amd.h) ----

dedisp_error dedisp_amd_set_profiling(dedisp_plan plan, int enable)
{
    if (!plan) return DEDISP_INVALID_PLAN;
    dedisp::GPUPlan* g = as_gpu(plan);
    if (!g) return DEDISP_INVALID_PLAN;
    g->set_profiling(enable != 0);
    return DEDISP_NO_ERROR;
}

dedisp_error dedisp_amd_get_profile(const dedisp_plan plan, dedisp_amd_profile* out)
{
    if (!plan) return DEDISP_INVALID_PLAN;
    if (!out) return DEDISP_INVALID_POINTER;
    dedisp::GPUPlan* g = as_gpu(plan);
    if (!g) return DEDISP_INVALID_PLAN;
    *out = g->get_profile();
    return DEDISP_NO_ERROR;
}

dedisp_error dedisp_amd_fdd_set_output(dedisp_plan plan, dedisp_amd_fdd_output mode)
{
    if (!plan) return DEDISP_INVALID_PLAN;
    dedisp::FDDGPUPlan* f = as_fdd(plan);
    if (!f) return DEDISP_INVALID_PLAN;
    return guarded([&] { f->set_output_mode(mode); });
}

dedisp_error dedisp_amd_fdd_set_precision(dedisp_plan plan, dedisp_amd_fdd_precision precision)
{
    if (!plan) return DEDISP_INVALID_PLAN;
    dedisp::FDDGPUPlan* f = as_fdd(plan);
    if (!f) return DEDISP_INVALID_PLAN;
    return guarded([&] { f->set_precision(precision); });
}

dedisp_error dedisp_amd_fdd_set_segmented(dedisp_plan plan, int enable, dedisp_size nfft)
{
    if (!plan) return DEDISP_INVALID_PLAN;
    dedisp::FDDGPUPlan* f = as_fdd(plan);
    if (!f) return DEDISP_INVALID_PLAN;
    return guarded([&] { f->set_segmented(enable, nfft); });
}

dedisp_size dedisp_amd_fdd_output_length(const dedisp_plan plan, dedisp_size nsamps)
{
    if (!plan) return 0;
    dedisp::FDDGPUPlan* f = as_fdd(plan);
    if (!f) return 0;
    return f->output_length(nsamps);
}

dedisp_error dedisp_amd_get_implementation(const dedisp_plan plan, dedisp_implementation* impl)
{
    if (!plan) return DEDISP_INVALID_PLAN;
    if (!impl) return DEDISP_INVALID_POINTER;
    *impl = plan->impl;
    return DEDISP_NO_ERROR;
}

dedisp_error dedisp_amd_generate_delay_table(dedisp_size nchans, dedisp_float dt, dedisp_float f0,
                                             dedisp_float df, dedisp_float* out)
{
    if (!out) return DEDISP_INVALID_POINTER;
    dedisp::make_delay_table(nchans, dt, f0, -std::abs(df), out);
    return DEDISP_NO_ERROR;
}

const char* dedisp_amd_version(void) { return "dedisp-mi355x 0.1.0 (gfx950)"; }

}  // extern "C"
