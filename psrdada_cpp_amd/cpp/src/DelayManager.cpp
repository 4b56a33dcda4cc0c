#include "psrdada_cpp/meerkat/fbfuse/DelayManager.hpp"

#include <errno.h>
#include <fcntl.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>

#include <cstring>
#include <stdexcept>
#include <string>

namespace psrdada_cpp {
namespace meerkat {
namespace fbfuse {

namespace {
std::runtime_error sys_error(std::string const& what) { return std::runtime_error(what + ": " + std::strerror(errno)); }
}  // namespace

DelayManager::DelayManager(PipelineConfig const& config, StreamType stream)
    : _config(config),
      _copy_stream(stream),
      _model_bytes(delay_model_bytes(config.cb_nbeams(), config.cb_nantennas())),
      _delay_model(nullptr),
      _delay_buffer_fd(-1),
      _delay_mutex_sem(SEM_FAILED),
      _delay_count_sem(SEM_FAILED),
      _last_sem_value(0),
      _updates(0)
{
    // DelayManager.cu:13-84
    _delay_buffer_fd = shm_open(_config.delay_buffer_shm().c_str(), O_RDONLY, 0);
    if (_delay_buffer_fd == -1) throw sys_error("Failed to open delay buffer shared memory");
    _delay_mutex_sem = sem_open(_config.delay_buffer_mutex().c_str(), 0);
    if (_delay_mutex_sem == SEM_FAILED) {
        close(_delay_buffer_fd);
        throw sys_error("Failed to open delay buffer mutex semaphore");
    }
    _delay_count_sem = sem_open(_config.delay_buffer_sem().c_str(), 0);
    if (_delay_count_sem == SEM_FAILED) {
        sem_close(_delay_mutex_sem);
        close(_delay_buffer_fd);
        throw sys_error("Failed to open delay buffer counting semaphore");
    }
    struct stat info;
    if (fstat(_delay_buffer_fd, &info) == -1) throw sys_error("Could not fstat the delay buffer shared memory");
    if ((std::size_t)info.st_size != _model_bytes)
        throw std::runtime_error("Shared memory buffer had unexpected size: " + std::to_string(info.st_size));
    _delay_model = mmap(nullptr, _model_bytes, PROT_READ, MAP_SHARED, _delay_buffer_fd, 0);
    if (_delay_model == MAP_FAILED) throw sys_error("MMAP on delay model buffer failed");
    _delays.resize(_config.cb_nbeams() * _config.cb_nantennas());
}

DelayManager::~DelayManager()
{
    if (_delay_model && _delay_model != MAP_FAILED) munmap(_delay_model, _model_bytes);
    if (_delay_buffer_fd != -1) close(_delay_buffer_fd);
    if (_delay_count_sem != SEM_FAILED) sem_close(_delay_count_sem);
    if (_delay_mutex_sem != SEM_FAILED) sem_close(_delay_mutex_sem);
}

bool DelayManager::update_available()
{
    int count = 0;
    if (sem_getvalue(_delay_count_sem, &count) != 0) throw sys_error("Unable to retrieve value of counting semaphore");
    if (count == _last_sem_value) return false;
    _last_sem_value = count;
    return true;
}

DelayManager::DelayVectorType const& DelayManager::delays()
{
    if (update_available()) {
        if (sem_wait(_delay_mutex_sem) != 0) throw sys_error("Unable to wait on mutex semaphore");
        // An async copy on the dedicated stream, then wait on the host only.
        const int st = fbf_memcpy_h2d(_delays.data(), delay_model_delays(_delay_model), _delays.bytes(), _copy_stream);
        const int st2 = st == FBF_OK ? fbf_stream_synchronize(_copy_stream) : st;
        if (sem_post(_delay_mutex_sem) != 0) throw sys_error("Unable to release mutex semaphore");
        FBF_ERROR_CHECK(st2);
        ++_updates;
    }
    return _delays;
}

double DelayManager::epoch() const { return static_cast<DelayModelHeader const*>(_delay_model)->epoch; }

double DelayManager::duration() const { return static_cast<DelayModelHeader const*>(_delay_model)->duration; }

DelayEngineSimulator::DelayEngineSimulator(PipelineConfig const& config)
    : _shm(config.delay_buffer_shm()),
      _sem(config.delay_buffer_sem()),
      _mutex(config.delay_buffer_mutex()),
      _bytes(delay_model_bytes(config.cb_nbeams(), config.cb_nantennas())),
      _fd(-1),
      _ptr(nullptr),
      _sem_id(SEM_FAILED),
      _mutex_id(SEM_FAILED)
{
    // DelayEngineSimulator.cu:14-67
    _fd = shm_open(_shm.c_str(), O_CREAT | O_RDWR, 0666);
    if (_fd == -1) throw sys_error("Failed to open shared memory named " + _shm);
    if (ftruncate(_fd, (off_t)_bytes) == -1) throw sys_error("Failed to ftruncate shared memory named " + _shm);
    _ptr = mmap(nullptr, _bytes, PROT_READ | PROT_WRITE, MAP_SHARED, _fd, 0);
    if (_ptr == MAP_FAILED) throw sys_error("Failed to mmap shared memory named " + _shm);
    std::memset(_ptr, 0, _bytes);
    sem_unlink(_sem.c_str());
    sem_unlink(_mutex.c_str());
    _sem_id = sem_open(_sem.c_str(), O_CREAT, 0666, 0);
    if (_sem_id == SEM_FAILED) throw sys_error("Failed to open delay buffer semaphore " + _sem);
    _mutex_id = sem_open(_mutex.c_str(), O_CREAT, 0666, 0);
    if (_mutex_id == SEM_FAILED) throw sys_error("Failed to open delay buffer mutex " + _mutex);
    sem_post(_mutex_id);   // the mutex starts available
}

DelayEngineSimulator::~DelayEngineSimulator()
{
    if (_ptr && _ptr != MAP_FAILED) munmap(_ptr, _bytes);
    if (_fd != -1) close(_fd);
    shm_unlink(_shm.c_str());
    if (_sem_id != SEM_FAILED) sem_close(_sem_id);
    if (_mutex_id != SEM_FAILED) sem_close(_mutex_id);
    sem_unlink(_sem.c_str());
    sem_unlink(_mutex.c_str());
}

DelayModelHeader* DelayEngineSimulator::header() { return static_cast<DelayModelHeader*>(_ptr); }

float2* DelayEngineSimulator::delays() { return delay_model_delays(_ptr); }

void DelayEngineSimulator::update_delays() { sem_post(_sem_id); }

}  // namespace fbfuse
}  // namespace meerkat
}  // namespace psrdada_cpp
