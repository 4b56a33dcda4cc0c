// Ring buffer of header and data blocks with DADA semantics, shared by the
// write and read clients.
//
// libpsrdada (ipcbuf rings) is not available, so this is a project-written
// ring that keeps the same client contract (dada_write_client.hpp,
// dada_read_client.hpp, dada_db.hpp): a ring is found by key_t, a writer
// fills blocks with next()/release(), one reader consumes them in order, a
// released block carries its used byte count, release(true) marks end of
// data, and a new reader attaching starts the next observation.
//
// Two placements:
//  * shared (the default, as dada_db.cpp:95-110 creates SysV segments): the
//    data ring lives in the SysV shared-memory segment `key`, the header
//    ring in `key + 1`, so a writer or reader in any process attaches by key
//    (an F-engine capture process writing, the beamformer reading).  Each
//    segment holds a control block with a process-shared, robust mutex and
//    condition variable, the per-block used counts and end-of-data flags,
//    then the blocks.  The blocks are ordinary pageable memory: the
//    reference pins them for device copies (dada_cuda_dbregister,
//    dada_client_base.cpp:78-88), which needs hipHostRegister.
//  * process: the same control blocks in this process's memory with blocks
//    in pinned host memory, found through a per-process key registry, for
//    rings whose writer and reader are threads of one process.
#ifndef PSRDADA_CPP_DADA_RING_HPP
#define PSRDADA_CPP_DADA_RING_HPP

#include <pthread.h>
#include <sys/types.h>

#include <cstddef>
#include <cstdint>
#include <memory>
#include <vector>

namespace psrdada_cpp {

// One sub-ring (data or header): its control block and its blocks, wherever
// they live.
class DadaSubRing {
public:
    // State shared by every process attached to the sub-ring.
    struct Control {
        uint64_t magic;
        uint64_t nbufs, bufsz, stride;
        uint64_t wr, rd, full;
        uint32_t eod_seen, writer_locked;
        int32_t writer_pid;
        uint32_t reserved;
        pthread_mutex_t mutex;   // process-shared, robust
        pthread_cond_t cond;     // process-shared
    };
    // Bytes of the control block and the per-block flags (page rounded), and
    // of a whole region: control, flags, then the blocks.
    static std::size_t control_bytes(std::size_t nbufs);
    static std::size_t bytes_needed(std::size_t nbufs, std::size_t bufsz);
    // Lays out a fresh control block and flags at `mem` (control_bytes).
    static void initialise(void* mem, std::size_t nbufs, std::size_t bufsz);
    DadaSubRing() = default;
    // A view of an initialised control region whose blocks follow it
    // (`blocks` empty) or are the given pointers.
    void attach(void* mem, std::vector<char*> blocks = {});

    std::size_t nbufs() const { return (std::size_t)_ctl->nbufs; }
    std::size_t bufsz() const { return (std::size_t)_ctl->bufsz; }
    char* acquire_write(std::size_t& index);
    void commit_write(std::size_t index, std::size_t used, bool eod);
    char* acquire_read(std::size_t& index, std::size_t& used, bool& eod);
    void commit_read(std::size_t index);
    bool at_end();
    void reset_eod();
    bool try_lock_writer();
    void unlock_writer();

private:
    class Lock;
    Control* _ctl = nullptr;
    uint64_t* _used = nullptr;
    uint8_t* _eod = nullptr;
    std::vector<char*> _blocks;
};

class DadaRing {
public:
    enum class Placement { shared, process };

    // A ring of this process (Placement::process) with pinned blocks.
    DadaRing(std::size_t nbufs, std::size_t bufsz, std::size_t nhdrs, std::size_t hdrsz);
    ~DadaRing();
    DadaRing(DadaRing const&) = delete;

    std::size_t data_buffer_size() const { return _data.bufsz(); }
    std::size_t header_buffer_size() const { return _hdr.bufsz(); }
    std::size_t data_buffer_count() const { return _data.nbufs(); }
    std::size_t header_buffer_count() const { return _hdr.nbufs(); }
    Placement placement() const { return _placement; }

    // Write side: acquire_write blocks until a slot is free; commit_write
    // publishes it.  Read side: acquire_read blocks until a published slot
    // (or end of data) is available; commit_read frees it.
    char* acquire_write(bool header, std::size_t& index) { return sub(header).acquire_write(index); }
    void commit_write(bool header, std::size_t index, std::size_t used, bool eod)
    {
        sub(header).commit_write(index, used, eod);
    }
    char* acquire_read(bool header, std::size_t& index, std::size_t& used, bool& eod)
    {
        return sub(header).acquire_read(index, used, eod);
    }
    void commit_read(bool header, std::size_t index) { sub(header).commit_read(index); }
    bool at_end(bool header) { return sub(header).at_end(); }
    // Clears end-of-data so that a new observation can follow.
    void reset_eod()
    {
        _data.reset_eod();
        _hdr.reset_eod();
    }
    // Lock for the single writer (dada_write_client.cpp: lock/release); a
    // lock held by a process that has exited is taken over.
    bool try_lock_writer() { return _data.try_lock_writer(); }
    void unlock_writer() { _data.unlock_writer(); }

    // Rings of this process, found by key before the shared segments.
    static void register_ring(key_t key, std::shared_ptr<DadaRing> ring);
    static void unregister_ring(key_t key);
    // The ring with this key: this process's, else the shared segments
    // `key` (data) and `key + 1` (header), attached for the caller.
    static std::shared_ptr<DadaRing> find(key_t key);

    // Shared segments: create (exclusive) and remove.
    static void create_shared(key_t key, std::size_t nbufs, std::size_t bufsz, std::size_t nhdrs, std::size_t hdrsz);
    static void remove_shared(key_t key);
    static bool shared_exists(key_t key);

private:
    struct SharedTag {};
    DadaRing(SharedTag, key_t key);
    DadaSubRing& sub(bool header) { return header ? _hdr : _data; }
    Placement _placement;
    DadaSubRing _data, _hdr;
    // process placement: the memory this object owns
    std::vector<char> _data_ctl, _hdr_ctl;
    std::vector<char*> _pinned;
    // shared placement: the attached segments
    void* _data_seg = nullptr;
    void* _hdr_seg = nullptr;
};

// Creates and owns a ring (dada_db.hpp:20-96) under a fresh key 0xdadaXXXX
// (even, with key + 1 free too, as dada_db.cpp:49-62 picks them).
class DadaDB {
public:
    DadaDB(std::size_t nbufs = 8, std::size_t bufsz = 4096, std::size_t nhdrs = 4, std::size_t hdrsz = 4096,
           DadaRing::Placement placement = DadaRing::Placement::shared);
    ~DadaDB();
    DadaDB(DadaDB const&) = delete;
    void create();
    void destroy();
    key_t key() const { return _key; }
    std::size_t data_buffer_size() const { return _bufsz; }
    std::size_t header_buffer_size() const { return _hdrsz; }
    std::size_t num_data_buffers() const { return _nbufs; }
    std::size_t num_header_buffers() const { return _nhdrs; }

private:
    std::size_t _nbufs, _bufsz, _nhdrs, _hdrsz;
    DadaRing::Placement _placement;
    key_t _key;
    bool _created;
};

}  // namespace psrdada_cpp

#endif
