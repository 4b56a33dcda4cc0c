// PipelineConfig (PipelineConfig.cpp:9-195): defaults are the reference's
// compile-time constants and runtime defaults.
#include "psrdada_cpp/meerkat/fbfuse/PipelineConfig.hpp"

namespace psrdada_cpp {
namespace meerkat {
namespace fbfuse {

PipelineConfig::PipelineConfig()
    : _delay_buffer_shm("fbfuse_delays_shm"),
      _delay_buffer_mutex("fbfuse_delays_mutex"),
      _delay_buffer_sem("fbfuse_buffer_counter"),
      _input_dada_key(0xdada),
      _cb_dada_key(0xcaca),
      _ib_dada_key(0xeaea),
      _cfreq(0.0f),
      _bw(0.0f),
      _channel_frequencies_stale(true),
      _input_level(32.0f),
      _output_level(24.0f),
      _cb_power_scaling(0.0f),
      _cb_power_offset(0.0f),
      _ib_power_scaling(0.0f),
      _ib_power_offset(0.0f),
      _quant_mode(FBF_QUANT_TRUNC),
      _total_nantennas(64),
      _nchans(64),
      _total_nchans(4096),
      _nsamples_per_heap(256),
      _cb_tscrunch(16),
      _cb_fscrunch(4),
      _cb_nantennas(44),
      _cb_antenna_offset(0),
      _cb_nbeams(1024),
      _cb_heap_size(8192),
      _ib_tscrunch(16),
      _ib_fscrunch(2)
{
    update_power_offsets_and_scalings();
}

PipelineConfig::~PipelineConfig() {}

std::string const& PipelineConfig::delay_buffer_shm() const { return _delay_buffer_shm; }
void PipelineConfig::delay_buffer_shm(std::string const& key) { _delay_buffer_shm = key; }
std::string const& PipelineConfig::delay_buffer_mutex() const { return _delay_buffer_mutex; }
void PipelineConfig::delay_buffer_mutex(std::string const& key) { _delay_buffer_mutex = key; }
std::string const& PipelineConfig::delay_buffer_sem() const { return _delay_buffer_sem; }
void PipelineConfig::delay_buffer_sem(std::string const& key) { _delay_buffer_sem = key; }
key_t PipelineConfig::input_dada_key() const { return _input_dada_key; }
void PipelineConfig::input_dada_key(key_t key) { _input_dada_key = key; }
key_t PipelineConfig::cb_dada_key() const { return _cb_dada_key; }
void PipelineConfig::cb_dada_key(key_t key) { _cb_dada_key = key; }
key_t PipelineConfig::ib_dada_key() const { return _ib_dada_key; }
void PipelineConfig::ib_dada_key(key_t key) { _ib_dada_key = key; }

float PipelineConfig::centre_frequency() const { return _cfreq; }
void PipelineConfig::centre_frequency(float cfreq)
{
    _cfreq = cfreq;
    _channel_frequencies_stale = true;
}
float PipelineConfig::bandwidth() const { return _bw; }
void PipelineConfig::bandwidth(float bw)
{
    _bw = bw;
    _channel_frequencies_stale = true;
}

std::vector<float> const& PipelineConfig::channel_frequencies() const
{
    if (_channel_frequencies_stale) calculate_channel_frequencies();
    return _channel_frequencies;
}

void PipelineConfig::calculate_channel_frequencies() const
{
    _channel_frequencies.assign(_nchans, 0.0f);
    FBF_ERROR_CHECK(fbf_channel_frequencies(_cfreq, _bw, (int)_nchans, _channel_frequencies.data()));
    _channel_frequencies_stale = false;
}

std::vector<double> PipelineConfig::channel_frequencies_f64() const
{
    std::vector<double> f(_nchans);
    FBF_ERROR_CHECK(fbf_channel_frequencies_f64((double)_cfreq, (double)_bw, (int)_nchans, f.data()));
    return f;
}

void PipelineConfig::output_level(float level)
{
    _output_level = level;
    update_power_offsets_and_scalings();
}
float PipelineConfig::output_level() const { return _output_level; }
void PipelineConfig::input_level(float level)
{
    _input_level = level;
    update_power_offsets_and_scalings();
}
float PipelineConfig::input_level() const { return _input_level; }

void PipelineConfig::update_power_offsets_and_scalings()
{
    FBF_ERROR_CHECK(fbf_power_levels(_input_level, _output_level, (int)_cb_nantennas, (int)_cb_tscrunch,
                                     (int)_cb_fscrunch, (int)_total_nantennas, (int)_ib_tscrunch,
                                     (int)_ib_fscrunch, (int)npol(), &_cb_power_offset, &_cb_power_scaling,
                                     &_ib_power_offset, &_ib_power_scaling));
}

float PipelineConfig::cb_power_scaling() const { return _cb_power_scaling; }
float PipelineConfig::cb_power_offset() const { return _cb_power_offset; }
float PipelineConfig::ib_power_scaling() const { return _ib_power_scaling; }
float PipelineConfig::ib_power_offset() const { return _ib_power_offset; }

static std::vector<float> per_channel(std::vector<float> const& explicit_levels, std::size_t n, float scalar)
{
    if (explicit_levels.size() == n) return explicit_levels;
    return std::vector<float>(n, scalar);
}

std::vector<float> PipelineConfig::cb_power_offsets() const
{
    return per_channel(_cb_offsets, _nchans / _cb_fscrunch, _cb_power_offset);
}
std::vector<float> PipelineConfig::cb_power_scalings() const
{
    return per_channel(_cb_scalings, _nchans / _cb_fscrunch, _cb_power_scaling);
}
std::vector<float> PipelineConfig::ib_power_offsets() const
{
    return per_channel(_ib_offsets, _nchans / _ib_fscrunch, _ib_power_offset);
}
std::vector<float> PipelineConfig::ib_power_scalings() const
{
    return per_channel(_ib_scalings, _nchans / _ib_fscrunch, _ib_power_scaling);
}

void PipelineConfig::cb_power_levels(std::vector<float> const& offsets, std::vector<float> const& scalings)
{
    if (offsets.size() != scalings.size()) throw std::runtime_error("cb_power_levels: size mismatch");
    _cb_offsets = offsets;
    _cb_scalings = scalings;
}

void PipelineConfig::ib_power_levels(std::vector<float> const& offsets, std::vector<float> const& scalings)
{
    if (offsets.size() != scalings.size()) throw std::runtime_error("ib_power_levels: size mismatch");
    _ib_offsets = offsets;
    _ib_scalings = scalings;
}

int PipelineConfig::quantisation_mode() const { return _quant_mode; }
void PipelineConfig::quantisation_mode(int mode)
{
    if (mode != FBF_QUANT_ROUND && mode != FBF_QUANT_TRUNC) throw std::runtime_error("bad quantisation mode");
    _quant_mode = mode;
}

#define FBF_GEOMETRY_SETTER(name, member)              \
    void PipelineConfig::name(std::size_t v)           \
    {                                                  \
        if (v == 0) throw std::runtime_error(#name " must be positive"); \
        member = v;                                    \
        _channel_frequencies_stale = true;             \
        update_power_offsets_and_scalings();           \
    }

FBF_GEOMETRY_SETTER(cb_tscrunch, _cb_tscrunch)
FBF_GEOMETRY_SETTER(cb_fscrunch, _cb_fscrunch)
FBF_GEOMETRY_SETTER(cb_nantennas, _cb_nantennas)
FBF_GEOMETRY_SETTER(cb_nbeams, _cb_nbeams)
FBF_GEOMETRY_SETTER(cb_heap_size, _cb_heap_size)
FBF_GEOMETRY_SETTER(ib_tscrunch, _ib_tscrunch)
FBF_GEOMETRY_SETTER(ib_fscrunch, _ib_fscrunch)
FBF_GEOMETRY_SETTER(total_nantennas, _total_nantennas)
FBF_GEOMETRY_SETTER(nchans, _nchans)
FBF_GEOMETRY_SETTER(total_nchans, _total_nchans)
FBF_GEOMETRY_SETTER(nsamples_per_heap, _nsamples_per_heap)

void PipelineConfig::cb_antenna_offset(std::size_t v) { _cb_antenna_offset = v; }

}  // namespace fbfuse
}  // namespace meerkat
}  // namespace psrdada_cpp
