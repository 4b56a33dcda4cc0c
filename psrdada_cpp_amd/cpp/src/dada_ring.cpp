#include "psrdada_cpp/dada_ring.hpp"

#include <atomic>
#include <map>
#include <stdexcept>
#include <string>

#include "psrdada_cpp/common.hpp"

namespace psrdada_cpp {

namespace {
std::mutex g_registry_mutex;
std::map<key_t, std::shared_ptr<DadaRing>>& registry()
{
    static std::map<key_t, std::shared_ptr<DadaRing>> r;
    return r;
}
std::atomic<unsigned> g_next_key{1};

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
}  // namespace

DadaRing::DadaRing(std::size_t nbufs, std::size_t bufsz, std::size_t nhdrs, std::size_t hdrsz)
    : _nbufs(nbufs), _bufsz(bufsz), _nhdrs(nhdrs), _hdrsz(hdrsz)
{
    if (!nbufs || !bufsz || !nhdrs || !hdrsz) throw std::runtime_error("DadaRing: sizes must be positive");
    for (auto* s : {&_data, &_hdr}) {
        const bool h = s == &_hdr;
        const std::size_t n = h ? nhdrs : nbufs, sz = h ? hdrsz : bufsz;
        for (std::size_t i = 0; i < n; ++i) s->blocks.push_back(pinned_block(sz));
        s->used.assign(n, 0);
        s->eod.assign(n, false);
    }
}

DadaRing::~DadaRing()
{
    for (auto* s : {&_data, &_hdr})
        for (char* b : s->blocks)
            if (fbf_host_free(b) != FBF_OK) ::operator delete(b);
}

char* DadaRing::acquire_write(bool header, std::size_t& index)
{
    std::unique_lock<std::mutex> lock(_mutex);
    Sub& s = sub(header);
    const std::size_t n = s.blocks.size();
    _cv.wait(lock, [&] { return s.full < n; });
    index = s.wr;
    return s.blocks[index];
}

void DadaRing::commit_write(bool header, std::size_t index, std::size_t used, bool eod)
{
    std::lock_guard<std::mutex> lock(_mutex);
    Sub& s = sub(header);
    if (index != s.wr) throw std::runtime_error("DadaRing: blocks must be released in order");
    s.used[index] = used;
    s.eod[index] = eod;
    s.wr = (s.wr + 1) % s.blocks.size();
    ++s.full;
    _cv.notify_all();
}

char* DadaRing::acquire_read(bool header, std::size_t& index, std::size_t& used, bool& eod)
{
    std::unique_lock<std::mutex> lock(_mutex);
    Sub& s = sub(header);
    _cv.wait(lock, [&] { return s.full > 0; });
    index = s.rd;
    used = s.used[index];
    eod = s.eod[index];
    return s.blocks[index];
}

void DadaRing::commit_read(bool header, std::size_t index)
{
    std::lock_guard<std::mutex> lock(_mutex);
    Sub& s = sub(header);
    if (index != s.rd) throw std::runtime_error("DadaRing: blocks must be released in order");
    if (s.eod[index]) s.eod_seen = true;
    s.rd = (s.rd + 1) % s.blocks.size();
    --s.full;
    _cv.notify_all();
}

bool DadaRing::at_end(bool header)
{
    std::lock_guard<std::mutex> lock(_mutex);
    return sub(header).eod_seen;
}

void DadaRing::reset_eod()
{
    std::lock_guard<std::mutex> lock(_mutex);
    _data.eod_seen = _hdr.eod_seen = false;
}

bool DadaRing::try_lock_writer()
{
    std::lock_guard<std::mutex> lock(_mutex);
    if (_writer_locked) return false;
    _writer_locked = true;
    return true;
}

void DadaRing::unlock_writer()
{
    std::lock_guard<std::mutex> lock(_mutex);
    _writer_locked = false;
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
    std::lock_guard<std::mutex> lock(g_registry_mutex);
    auto it = registry().find(key);
    if (it == registry().end())
        throw std::runtime_error("No DADA ring with key " + std::to_string((long)key));
    return it->second;
}

DadaDB::DadaDB(std::size_t nbufs, std::size_t bufsz, std::size_t nhdrs, std::size_t hdrsz)
    : _nbufs(nbufs), _bufsz(bufsz), _nhdrs(nhdrs), _hdrsz(hdrsz), _key(0), _created(false)
{
}

DadaDB::~DadaDB() { destroy(); }

void DadaDB::create()
{
    if (_created) return;
    _key = (key_t)(0xdada0000u + (g_next_key++ & 0xffffu));
    DadaRing::register_ring(_key, std::make_shared<DadaRing>(_nbufs, _bufsz, _nhdrs, _hdrsz));
    _created = true;
}

void DadaDB::destroy()
{
    if (!_created) return;
    DadaRing::unregister_ring(_key);
    _created = false;
}

}  // namespace psrdada_cpp
