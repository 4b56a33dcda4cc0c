// FBFUSE pipeline configuration (PipelineConfig.hpp:14-336).
//
// Same getters and setters as the reference.  The geometry that the
// reference fixes at compile time (fbfuse_constants.hpp:4-39) is runtime
// state here, with the reference's values as defaults, so that one process
// can run several subarray shapes.  Power offsets and scalings follow
// PipelineConfig.cpp:109-127; per-output-channel vectors are available in
// addition to the reference's scalars (all channels equal by default).
#ifndef PSRDADA_CPP_MEERKAT_FBFUSE_PIPELINECONFIG_HPP
#define PSRDADA_CPP_MEERKAT_FBFUSE_PIPELINECONFIG_HPP

#include <sys/types.h>

#include <string>
#include <vector>

#include "psrdada_cpp/common.hpp"

namespace psrdada_cpp {
namespace meerkat {
namespace fbfuse {

class PipelineConfig {
public:
    PipelineConfig();
    ~PipelineConfig();
    PipelineConfig(PipelineConfig const&) = delete;

    // --- control-plane names and DADA keys (PipelineConfig.hpp:25-105) ---
    std::string const& delay_buffer_shm() const;
    void delay_buffer_shm(std::string const&);
    std::string const& delay_buffer_mutex() const;
    void delay_buffer_mutex(std::string const&);
    std::string const& delay_buffer_sem() const;
    void delay_buffer_sem(std::string const&);
    key_t input_dada_key() const;
    void input_dada_key(key_t);
    key_t cb_dada_key() const;
    void cb_dada_key(key_t);
    key_t ib_dada_key() const;
    void ib_dada_key(key_t);

    // --- sky frequencies (PipelineConfig.hpp:111-135) ---
    float centre_frequency() const;
    void centre_frequency(float cfreq);
    float bandwidth() const;
    void bandwidth(float bw);
    std::vector<float> const& channel_frequencies() const;
    // The same channel centres computed in double precision (the weights
    // generator forms its phase argument in FP64).
    std::vector<double> channel_frequencies_f64() const;

    // --- levels and quantisation (PipelineConfig.hpp:180-240) ---
    void output_level(float level);
    float output_level() const;
    void input_level(float level);
    float input_level() const;
    float cb_power_scaling() const;
    float cb_power_offset() const;
    float ib_power_scaling() const;
    float ib_power_offset() const;
    // Per output channel (F / fscrunch entries); the reference's scalar in
    // every channel unless set explicitly.
    std::vector<float> cb_power_offsets() const;
    std::vector<float> cb_power_scalings() const;
    std::vector<float> ib_power_offsets() const;
    std::vector<float> ib_power_scalings() const;
    void cb_power_levels(std::vector<float> const& offsets, std::vector<float> const& scalings);
    void ib_power_levels(std::vector<float> const& offsets, std::vector<float> const& scalings);
    // FBF_QUANT_TRUNC reproduces the reference cast (CoherentBeamformer.cu:169);
    // FBF_QUANT_ROUND rounds half away from zero.  Both saturate.
    int quantisation_mode() const;
    void quantisation_mode(int mode);

    // --- geometry (fbfuse_constants.hpp:4-39), runtime here ---
    std::size_t cb_tscrunch() const { return _cb_tscrunch; }
    std::size_t cb_fscrunch() const { return _cb_fscrunch; }
    std::size_t cb_nantennas() const { return _cb_nantennas; }
    std::size_t cb_antenna_offset() const { return _cb_antenna_offset; }
    std::size_t cb_nbeams() const { return _cb_nbeams; }
    std::size_t cb_nsamples_per_block() const { return _cb_tscrunch * 32; }
    std::size_t cb_heap_size() const { return _cb_heap_size; }
    std::size_t cb_nsamples_per_heap() const { return _cb_heap_size / (nchans() / cb_fscrunch()); }
    std::size_t ib_tscrunch() const { return _ib_tscrunch; }
    std::size_t ib_fscrunch() const { return _ib_fscrunch; }
    std::size_t ib_nantennas() const { return _total_nantennas; }
    std::size_t ib_anntena_offset() const { return 0; }
    std::size_t ib_nbeams() const { return 1; }
    std::size_t total_nantennas() const { return _total_nantennas; }
    std::size_t nchans() const { return _nchans; }
    std::size_t total_nchans() const { return _total_nchans; }
    std::size_t npol() const { return 2; }
    std::size_t nsamples_per_heap() const { return _nsamples_per_heap; }

    void cb_tscrunch(std::size_t v);
    void cb_fscrunch(std::size_t v);
    void cb_nantennas(std::size_t v);
    void cb_antenna_offset(std::size_t v);
    void cb_nbeams(std::size_t v);
    void cb_heap_size(std::size_t bytes);
    void ib_tscrunch(std::size_t v);
    void ib_fscrunch(std::size_t v);
    void total_nantennas(std::size_t v);
    void nchans(std::size_t v);
    void total_nchans(std::size_t v);
    void nsamples_per_heap(std::size_t v);

private:
    void calculate_channel_frequencies() const;
    void update_power_offsets_and_scalings();

    std::string _delay_buffer_shm;
    std::string _delay_buffer_mutex;
    std::string _delay_buffer_sem;
    key_t _input_dada_key;
    key_t _cb_dada_key;
    key_t _ib_dada_key;
    float _cfreq;
    float _bw;
    mutable std::vector<float> _channel_frequencies;
    mutable bool _channel_frequencies_stale;
    float _input_level;
    float _output_level;
    float _cb_power_scaling;
    float _cb_power_offset;
    float _ib_power_scaling;
    float _ib_power_offset;
    std::vector<float> _cb_offsets, _cb_scalings, _ib_offsets, _ib_scalings;  // explicit per-channel levels
    int _quant_mode;
    std::size_t _total_nantennas, _nchans, _total_nchans, _nsamples_per_heap;
    std::size_t _cb_tscrunch, _cb_fscrunch, _cb_nantennas, _cb_antenna_offset, _cb_nbeams, _cb_heap_size;
    std::size_t _ib_tscrunch, _ib_fscrunch;
};

}  // namespace fbfuse
}  // namespace meerkat
}  // namespace psrdada_cpp

#endif
