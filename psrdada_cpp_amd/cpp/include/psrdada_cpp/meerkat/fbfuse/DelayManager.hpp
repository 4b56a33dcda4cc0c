// Delay-model ingest from the control system (DelayManager.cuh:18-74).
//
// Opens the POSIX shared memory holding the DelayModel and the two named
// semaphores (a mutex and an update counter).  delays() checks the counter
// and, when it moved, copies the new delays to the device under the mutex
// (DelayManager.cu:116-185).
#ifndef PSRDADA_CPP_MEERKAT_FBFUSE_DELAYMANAGER_HPP
#define PSRDADA_CPP_MEERKAT_FBFUSE_DELAYMANAGER_HPP

#include <semaphore.h>

#include "psrdada_cpp/device_vector.hpp"
#include "psrdada_cpp/meerkat/fbfuse/DelayModel.hpp"
#include "psrdada_cpp/meerkat/fbfuse/PipelineConfig.hpp"

namespace psrdada_cpp {
namespace meerkat {
namespace fbfuse {

class DelayManager {
public:
    typedef float2 DelayType;
    typedef DeviceVector<DelayType> DelayVectorType;
    typedef double TimeType;

    DelayManager(PipelineConfig const& config, StreamType stream);
    ~DelayManager();
    DelayManager(DelayManager const&) = delete;

    DelayVectorType const& delays();
    double epoch() const;
    double duration() const;
    // Number of delay-model updates copied to the device so far.
    std::size_t update_count() const { return _updates; }

private:
    bool update_available();

    PipelineConfig const& _config;
    StreamType _copy_stream;
    std::size_t _model_bytes;
    void* _delay_model;
    int _delay_buffer_fd;
    sem_t* _delay_mutex_sem;
    sem_t* _delay_count_sem;
    int _last_sem_value;
    std::size_t _updates;
    DelayVectorType _delays;
};

// Stand-in for the control system in tests (DelayEngineSimulator.cuh:12-44):
// creates the shared memory and semaphores, and bumps the counter on update.
class DelayEngineSimulator {
public:
    explicit DelayEngineSimulator(PipelineConfig const& config);
    ~DelayEngineSimulator();
    DelayEngineSimulator(DelayEngineSimulator const&) = delete;
    DelayModelHeader* header();
    float2* delays();
    void update_delays();

private:
    std::string _shm, _sem, _mutex;
    std::size_t _bytes;
    int _fd;
    void* _ptr;
    sem_t* _sem_id;
    sem_t* _mutex_id;
};

}  // namespace fbfuse
}  // namespace meerkat
}  // namespace psrdada_cpp

#endif
