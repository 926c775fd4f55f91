// FDDGPUPlan: host side of the Fourier-domain dedispersion path.
#include "internal.hpp"

namespace dedisp {

struct FDDState {
    dedisp_amd_fdd_output output = DEDISP_AMD_FDD_FOURIER;
    dedisp_amd_fdd_precision precision = DEDISP_AMD_FDD_FP32;
    int segmented = 0;
    dedisp_size nfft = 0;
};

FDDGPUPlan::FDDGPUPlan(size_type nchans, float_type dt, float_type f0, float_type df, int device_idx)
    : GPUPlan(nchans, dt, f0, df, device_idx), m_state(new FDDState)
{
    if (nchans > kFddMaxChans) DD_THROW(DEDISP_NCHANS_EXCEEDS_LIMIT);
}

FDDGPUPlan::~FDDGPUPlan() {}

void FDDGPUPlan::set_output_mode(dedisp_amd_fdd_output mode) { m_state->output = mode; }
dedisp_amd_fdd_output FDDGPUPlan::get_output_mode() const { return m_state->output; }
void FDDGPUPlan::set_precision(dedisp_amd_fdd_precision p) { m_state->precision = p; }
dedisp_amd_fdd_precision FDDGPUPlan::get_precision() const { return m_state->precision; }
void FDDGPUPlan::set_segmented(int enable, dedisp_size nfft) { m_state->segmented = enable; m_state->nfft = nfft; }

dedisp_size FDDGPUPlan::output_length(dedisp_size nsamps) const
{
    if (m_state->output == DEDISP_AMD_FDD_FOURIER && !m_state->segmented)
        return round_up(nsamps + m_max_delay, 16384);
    return nsamps - m_max_delay;
}

void FDDGPUPlan::execute(size_type, const byte_type*, size_type, byte_type*, size_type, unsigned)
{
    DD_THROW(DEDISP_UNKNOWN_ERROR);
}

}  // namespace dedisp
