// Ring buffer of header and data blocks with DADA semantics, shared by the
// write and read clients.
//
// libpsrdada (SysV shared-memory ipcbuf rings) is not available, so this is a
// project-written ring that keeps the same client contract
// (dada_write_client.hpp, dada_read_client.hpp, dada_db.hpp): rings are
// looked up by key_t, a writer fills blocks with next()/release(), a reader
// consumes them in order, a released block carries its used byte count, and
// release(true) marks end of data.  Rings live in one process (threads play
// the roles of the reference's processes); their blocks are pinned host
// memory so that device copies into and out of them run at full speed.
#ifndef PSRDADA_CPP_DADA_RING_HPP
#define PSRDADA_CPP_DADA_RING_HPP

#include <sys/types.h>

#include <condition_variable>
#include <cstddef>
#include <memory>
#include <mutex>
#include <vector>

namespace psrdada_cpp {

class DadaRing {
public:
    DadaRing(std::size_t nbufs, std::size_t bufsz, std::size_t nhdrs, std::size_t hdrsz);
    ~DadaRing();
    DadaRing(DadaRing const&) = delete;

    std::size_t data_buffer_size() const { return _bufsz; }
    std::size_t header_buffer_size() const { return _hdrsz; }
    std::size_t data_buffer_count() const { return _nbufs; }
    std::size_t header_buffer_count() const { return _nhdrs; }

    // Sub-ring of blocks.  Write side: acquire_write blocks until a slot is
    // free; commit_write publishes it.  Read side: acquire_read blocks until a
    // published slot (or end of data) is available; commit_read frees it.
    struct Sub {
        std::vector<char*> blocks;
        std::vector<std::size_t> used;
        std::vector<bool> eod;
        std::size_t wr = 0, rd = 0, full = 0;
        bool eod_seen = false;
    };
    char* acquire_write(bool header, std::size_t& index);
    void commit_write(bool header, std::size_t index, std::size_t used, bool eod);
    char* acquire_read(bool header, std::size_t& index, std::size_t& used, bool& eod);
    void commit_read(bool header, std::size_t index);
    bool at_end(bool header);
    // Clears end-of-data so that a new observation can follow.
    void reset_eod();
    // Lock for the single writer (dada_write_client.cpp: lock/release).
    bool try_lock_writer();
    void unlock_writer();

    // Key registry: rings are found by key like SysV rings.
    static void register_ring(key_t key, std::shared_ptr<DadaRing> ring);
    static void unregister_ring(key_t key);
    static std::shared_ptr<DadaRing> find(key_t key);

private:
    Sub& sub(bool header) { return header ? _hdr : _data; }
    std::size_t _nbufs, _bufsz, _nhdrs, _hdrsz;
    Sub _data, _hdr;
    std::mutex _mutex;
    std::condition_variable _cv;
    bool _writer_locked = false;
};

// Creates and owns a ring under a fresh key (dada_db.hpp:20-96; the key is
// 0xdada0000 + n like the reference's random test keys).
class DadaDB {
public:
    DadaDB(std::size_t nbufs = 8, std::size_t bufsz = 4096, std::size_t nhdrs = 4, std::size_t hdrsz = 4096);
    ~DadaDB();
    DadaDB(DadaDB const&) = delete;
    void create();
    void destroy();
    key_t key() const { return _key; }
    std::size_t data_buffer_size() const { return _bufsz; }
    std::size_t header_buffer_size() const { return _hdrsz; }

private:
    std::size_t _nbufs, _bufsz, _nhdrs, _hdrsz;
    key_t _key;
    bool _created;
};

}  // namespace psrdada_cpp

#endif
