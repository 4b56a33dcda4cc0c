#include "psrdada_cpp/dada_ring.hpp"

#include <errno.h>
#include <signal.h>
#include <sys/ipc.h>
#include <sys/shm.h>
#include <unistd.h>

#include <atomic>
#include <cstdlib>
#include <cstring>
#include <ctime>
#include <map>
#include <mutex>
#include <stdexcept>
#include <string>

#include "psrdada_cpp/common.hpp"

namespace psrdada_cpp {

namespace {

constexpr uint64_t kMagic = 0x3141444144464246ull;   // "FBFDADA1"
constexpr std::size_t kPage = 4096;

std::size_t round_up(std::size_t x, std::size_t a) { return (x + a - 1) / a * a; }

std::runtime_error sys_error(std::string const& what)
{
    return std::runtime_error(what + ": " + std::strerror(errno));
}

char* pinned_block(std::size_t bytes)
{
    void* p = nullptr;
    if (fbf_host_malloc(&p, bytes) != FBF_OK || !p) {
        // No device runtime available: plain host memory still works, copies
        // just run at pageable speed.
        p = ::operator new(bytes);
    }
    return static_cast<char*>(p);
}

void free_pinned(char* p)
{
    if (fbf_host_free(p) != FBF_OK) ::operator delete(p);
}

std::mutex g_registry_mutex;
std::map<key_t, std::shared_ptr<DadaRing>>& registry()
{
    static std::map<key_t, std::shared_ptr<DadaRing>> r;
    return r;
}

// A key that another process took between the check and the creation.
struct KeyInUse : std::runtime_error {
    KeyInUse() : std::runtime_error("DADA key in use") {}
};

// Whether the process that took a lock is gone.
bool process_gone(int32_t pid) { return pid > 0 && kill((pid_t)pid, 0) == -1 && errno == ESRCH; }

}  // namespace

// ---------------------------------------------------------------------------
// DadaSubRing
// ---------------------------------------------------------------------------

// The control block's mutex; a lock whose holder died (robust mutex) is
// recovered: the state it protects is only ever changed by single stores.
class DadaSubRing::Lock {
public:
    explicit Lock(Control* c) : _c(c)
    {
        const int rc = pthread_mutex_lock(&_c->mutex);
        if (rc == EOWNERDEAD)
            pthread_mutex_consistent(&_c->mutex);
        else if (rc != 0)
            throw std::runtime_error("DADA ring: mutex lock failed");
    }
    ~Lock() { pthread_mutex_unlock(&_c->mutex); }
    template <class Pred>
    void wait(Pred pred)
    {
        while (!pred()) {
            const int rc = pthread_cond_wait(&_c->cond, &_c->mutex);
            if (rc == EOWNERDEAD) pthread_mutex_consistent(&_c->mutex);
        }
    }
    void notify() { pthread_cond_broadcast(&_c->cond); }

private:
    Control* _c;
};

std::size_t DadaSubRing::control_bytes(std::size_t nbufs)
{
    return round_up(round_up(sizeof(Control), 64) + nbufs * sizeof(uint64_t) + nbufs, kPage);
}

std::size_t DadaSubRing::bytes_needed(std::size_t nbufs, std::size_t bufsz)
{
    return control_bytes(nbufs) + nbufs * round_up(bufsz, 64);
}

void DadaSubRing::initialise(void* mem, std::size_t nbufs, std::size_t bufsz)
{
    std::memset(mem, 0, control_bytes(nbufs));
    Control* c = static_cast<Control*>(mem);
    c->nbufs = nbufs;
    c->bufsz = bufsz;
    c->stride = round_up(bufsz, 64);
    pthread_mutexattr_t ma;
    pthread_mutexattr_init(&ma);
    pthread_mutexattr_setpshared(&ma, PTHREAD_PROCESS_SHARED);
    pthread_mutexattr_setrobust(&ma, PTHREAD_MUTEX_ROBUST);
    const int rm = pthread_mutex_init(&c->mutex, &ma);
    pthread_mutexattr_destroy(&ma);
    pthread_condattr_t ca;
    pthread_condattr_init(&ca);
    pthread_condattr_setpshared(&ca, PTHREAD_PROCESS_SHARED);
    pthread_condattr_setclock(&ca, CLOCK_MONOTONIC);
    const int rc = pthread_cond_init(&c->cond, &ca);
    pthread_condattr_destroy(&ca);
    if (rm != 0 || rc != 0) throw std::runtime_error("DADA ring: cannot initialise the process-shared lock");
    // Published last: an attaching process checks it.
    std::atomic_thread_fence(std::memory_order_release);
    c->magic = kMagic;
}

void DadaSubRing::attach(void* mem, std::vector<char*> blocks)
{
    _ctl = static_cast<Control*>(mem);
    if (_ctl->magic != kMagic) throw std::runtime_error("DADA ring: segment is not an initialised ring");
    char* flags = static_cast<char*>(mem) + round_up(sizeof(Control), 64);
    _used = reinterpret_cast<uint64_t*>(flags);
    _eod = reinterpret_cast<uint8_t*>(flags + _ctl->nbufs * sizeof(uint64_t));
    if (blocks.empty()) {
        char* base = static_cast<char*>(mem) + control_bytes(_ctl->nbufs);
        for (std::size_t i = 0; i < _ctl->nbufs; ++i) blocks.push_back(base + i * _ctl->stride);
    }
    if (blocks.size() != _ctl->nbufs) throw std::runtime_error("DADA ring: block count mismatch");
    _blocks = std::move(blocks);
}

char* DadaSubRing::acquire_write(std::size_t& index)
{
    Lock lock(_ctl);
    lock.wait([&] { return _ctl->full < _ctl->nbufs; });
    index = (std::size_t)_ctl->wr;
    return _blocks[index];
}

void DadaSubRing::commit_write(std::size_t index, std::size_t used, bool eod)
{
    Lock lock(_ctl);
    if (index != _ctl->wr) throw std::runtime_error("DADA ring: blocks must be released in order");
    _used[index] = used;
    _eod[index] = eod ? 1 : 0;
    _ctl->wr = (_ctl->wr + 1) % _ctl->nbufs;
    ++_ctl->full;
    lock.notify();
}

char* DadaSubRing::acquire_read(std::size_t& index, std::size_t& used, bool& eod)
{
    Lock lock(_ctl);
    lock.wait([&] { return _ctl->full > 0; });
    index = (std::size_t)_ctl->rd;
    used = (std::size_t)_used[index];
    eod = _eod[index] != 0;
    return _blocks[index];
}

void DadaSubRing::commit_read(std::size_t index)
{
    Lock lock(_ctl);
    if (index != _ctl->rd) throw std::runtime_error("DADA ring: blocks must be released in order");
    if (_eod[index]) _ctl->eod_seen = 1;
    _ctl->rd = (_ctl->rd + 1) % _ctl->nbufs;
    --_ctl->full;
    lock.notify();
}

bool DadaSubRing::at_end()
{
    Lock lock(_ctl);
    return _ctl->eod_seen != 0;
}

void DadaSubRing::reset_eod()
{
    Lock lock(_ctl);
    _ctl->eod_seen = 0;
}

bool DadaSubRing::try_lock_writer()
{
    Lock lock(_ctl);
    if (_ctl->writer_locked && !process_gone(_ctl->writer_pid)) return false;
    _ctl->writer_locked = 1;
    _ctl->writer_pid = (int32_t)getpid();
    return true;
}

void DadaSubRing::unlock_writer()
{
    Lock lock(_ctl);
    _ctl->writer_locked = 0;
    _ctl->writer_pid = 0;
}

// ---------------------------------------------------------------------------
// DadaRing
// ---------------------------------------------------------------------------

DadaRing::DadaRing(std::size_t nbufs, std::size_t bufsz, std::size_t nhdrs, std::size_t hdrsz)
    : _placement(Placement::process)
{
    if (!nbufs || !bufsz || !nhdrs || !hdrsz) throw std::runtime_error("DadaRing: sizes must be positive");
    _data_ctl.resize(DadaSubRing::control_bytes(nbufs) + 64);
    _hdr_ctl.resize(DadaSubRing::control_bytes(nhdrs) + 64);
    auto aligned = [](std::vector<char>& v) { return reinterpret_cast<void*>(round_up((uintptr_t)v.data(), 64)); };
    DadaSubRing::initialise(aligned(_data_ctl), nbufs, bufsz);
    DadaSubRing::initialise(aligned(_hdr_ctl), nhdrs, hdrsz);
    std::vector<char*> data_blocks, hdr_blocks;
    for (std::size_t i = 0; i < nbufs; ++i) data_blocks.push_back(pinned_block(bufsz));
    for (std::size_t i = 0; i < nhdrs; ++i) hdr_blocks.push_back(pinned_block(hdrsz));
    _pinned = data_blocks;
    _pinned.insert(_pinned.end(), hdr_blocks.begin(), hdr_blocks.end());
    _data.attach(aligned(_data_ctl), data_blocks);
    _hdr.attach(aligned(_hdr_ctl), hdr_blocks);
}

DadaRing::DadaRing(SharedTag, key_t key) : _placement(Placement::shared)
{
    auto attach_seg = [](key_t k) -> void* {
        const int id = shmget(k, 0, 0);
        if (id == -1) throw sys_error("No DADA ring with key " + std::to_string((long)k));
        void* p = shmat(id, nullptr, 0);
        if (p == (void*)-1) throw sys_error("Cannot attach the DADA ring segment with key " + std::to_string((long)k));
        return p;
    };
    _data_seg = attach_seg(key);
    try {
        _hdr_seg = attach_seg(key + 1);
        _data.attach(_data_seg);
        _hdr.attach(_hdr_seg);
    } catch (...) {
        shmdt(_data_seg);
        if (_hdr_seg) shmdt(_hdr_seg);
        throw;
    }
}

DadaRing::~DadaRing()
{
    if (_placement == Placement::process) {
        for (char* b : _pinned) free_pinned(b);
    } else {
        if (_data_seg) shmdt(_data_seg);
        if (_hdr_seg) shmdt(_hdr_seg);
    }
}

void DadaRing::register_ring(key_t key, std::shared_ptr<DadaRing> ring)
{
    std::lock_guard<std::mutex> lock(g_registry_mutex);
    if (registry().count(key)) throw std::runtime_error("DadaRing: key already in use");
    registry()[key] = std::move(ring);
}

void DadaRing::unregister_ring(key_t key)
{
    std::lock_guard<std::mutex> lock(g_registry_mutex);
    registry().erase(key);
}

std::shared_ptr<DadaRing> DadaRing::find(key_t key)
{
    {
        std::lock_guard<std::mutex> lock(g_registry_mutex);
        auto it = registry().find(key);
        if (it != registry().end()) return it->second;
    }
    return std::shared_ptr<DadaRing>(new DadaRing(SharedTag{}, key));
}

bool DadaRing::shared_exists(key_t key) { return shmget(key, 0, 0) != -1; }

void DadaRing::create_shared(key_t key, std::size_t nbufs, std::size_t bufsz, std::size_t nhdrs, std::size_t hdrsz)
{
    if (!nbufs || !bufsz || !nhdrs || !hdrsz) throw std::runtime_error("DadaRing: sizes must be positive");
    auto make = [](key_t k, std::size_t n, std::size_t sz) {
        const int id = shmget(k, DadaSubRing::bytes_needed(n, sz), IPC_CREAT | IPC_EXCL | 0666);
        if (id == -1 && errno == EEXIST) throw KeyInUse();
        if (id == -1) throw sys_error("Cannot create the DADA ring segment with key " + std::to_string((long)k));
        void* p = shmat(id, nullptr, 0);
        if (p == (void*)-1) {
            shmctl(id, IPC_RMID, nullptr);
            throw sys_error("Cannot attach the new DADA ring segment with key " + std::to_string((long)k));
        }
        DadaSubRing::initialise(p, n, sz);
        shmdt(p);
    };
    make(key, nbufs, bufsz);
    try {
        make(key + 1, nhdrs, hdrsz);
    } catch (...) {
        const int id = shmget(key, 0, 0);
        if (id != -1) shmctl(id, IPC_RMID, nullptr);
        throw;
    }
}

void DadaRing::remove_shared(key_t key)
{
    for (key_t k : {key, key + 1}) {
        const int id = shmget(k, 0, 0);
        if (id != -1) shmctl(id, IPC_RMID, nullptr);   // freed once every process has detached
    }
}

// ---------------------------------------------------------------------------
// DadaDB
// ---------------------------------------------------------------------------

DadaDB::DadaDB(std::size_t nbufs, std::size_t bufsz, std::size_t nhdrs, std::size_t hdrsz,
               DadaRing::Placement placement)
    : _nbufs(nbufs), _bufsz(bufsz), _nhdrs(nhdrs), _hdrsz(hdrsz), _placement(placement), _key(0), _created(false)
{
}

DadaDB::~DadaDB()
{
    try {
        destroy();
    } catch (...) {
    }
}

void DadaDB::create()
{
    if (_created) throw std::runtime_error("DADA data blocks already created");
    // An even key 0xdadaXXXX with key + 1 free as well, in this process's
    // registry and among the shared segments.
    unsigned seed = (unsigned)std::time(nullptr) ^ ((unsigned)getpid() << 8);
    for (int attempt = 0; attempt < 64; ++attempt) {
        seed = seed * 1103515245u + 12345u;
        const key_t k = (key_t)(0xdada0000u + (((seed >> 8) & 0xffffu) & ~1u));
        bool used = DadaRing::shared_exists(k) || DadaRing::shared_exists(k + 1);
        {
            std::lock_guard<std::mutex> lock(g_registry_mutex);
            used = used || registry().count(k) || registry().count(k + 1);
        }
        if (used) continue;
        if (_placement == DadaRing::Placement::shared) {
            try {
                DadaRing::create_shared(k, _nbufs, _bufsz, _nhdrs, _hdrsz);
            } catch (KeyInUse const&) {
                continue;   // taken meanwhile by another process
            }
        } else {
            DadaRing::register_ring(k, std::make_shared<DadaRing>(_nbufs, _bufsz, _nhdrs, _hdrsz));
        }
        _key = k;
        _created = true;
        return;
    }
    throw std::runtime_error("Unable to find a free DADA key");
}

void DadaDB::destroy()
{
    if (!_created) return;
    if (_placement == DadaRing::Placement::shared)
        DadaRing::remove_shared(_key);
    else
        DadaRing::unregister_ring(_key);
    _created = false;
}

}  // namespace psrdada_cpp
