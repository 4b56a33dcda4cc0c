// DADA write and read clients over DadaRing (dada_write_client.cpp:5-161,
// dada_read_client.cpp:5-197).
#include <stdexcept>

#include "psrdada_cpp/dada_read_client.hpp"
#include "psrdada_cpp/dada_write_client.hpp"

namespace psrdada_cpp {

DadaWriteClient::DadaWriteClient(key_t key, MultiLog& log)
    : _key(key), _log(log), _ring(DadaRing::find(key)), _header_stream(*this), _data_stream(*this)
{
    if (!_ring->try_lock_writer()) throw std::runtime_error("DadaWriteClient: ring already has a writer");
    _log.write(0, "write client locked the ring");
}

DadaWriteClient::~DadaWriteClient() { _ring->unlock_writer(); }

void DadaWriteClient::reset()
{
    _ring->unlock_writer();
    _ring = DadaRing::find(_key);
    if (!_ring->try_lock_writer()) throw std::runtime_error("DadaWriteClient: ring already has a writer");
}

RawBytes& DadaWriteClient::HeaderStream::next()
{
    if (_current_block) throw std::runtime_error("Previous header block not released");
    char* p = _parent._ring->acquire_write(true, _index);
    _current_block.reset(new RawBytes(p, _parent._ring->header_buffer_size(), 0));
    return *_current_block;
}

void DadaWriteClient::HeaderStream::release()
{
    if (!_current_block) throw std::runtime_error("No header block to release");
    _parent._ring->commit_write(true, _index, _current_block->used_bytes(), false);
    _current_block.reset();
}

RawBytes& DadaWriteClient::DataStream::next()
{
    if (_current_block) throw std::runtime_error("Previous data block not released");
    char* p = _parent._ring->acquire_write(false, _index);
    _current_block.reset(new RawBytes(p, _parent._ring->data_buffer_size(), 0));
    return *_current_block;
}

void DadaWriteClient::DataStream::release(bool eod)
{
    if (!_current_block) throw std::runtime_error("No data block to release");
    _parent._ring->commit_write(false, _index, _current_block->used_bytes(), eod);
    _current_block.reset();
    ++_block_idx;
}

DadaReadClient::DadaReadClient(key_t key, MultiLog& log)
    : _key(key), _log(log), _ring(DadaRing::find(key)), _header_stream(*this), _data_stream(*this)
{
    _ring->reset_eod();
}

DadaReadClient::~DadaReadClient() {}

RawBytes& DadaReadClient::HeaderStream::next()
{
    if (_current_block) throw std::runtime_error("Previous header block not released");
    std::size_t used = 0;
    bool eod = false;
    char* p = _parent._ring->acquire_read(true, _index, used, eod);
    _current_block.reset(new RawBytes(p, _parent._ring->header_buffer_size(), used));
    return *_current_block;
}

void DadaReadClient::HeaderStream::release()
{
    if (!_current_block) throw std::runtime_error("No header block to release");
    _parent._ring->commit_read(true, _index);
    _current_block.reset();
}

bool DadaReadClient::HeaderStream::at_end() const { return _parent._ring->at_end(true); }

RawBytes& DadaReadClient::DataStream::next()
{
    if (_current_block) throw std::runtime_error("Previous data block not released");
    std::size_t used = 0;
    bool eod = false;
    char* p = _parent._ring->acquire_read(false, _index, used, eod);
    _current_block.reset(new RawBytes(p, _parent._ring->data_buffer_size(), used));
    return *_current_block;
}

void DadaReadClient::DataStream::release()
{
    if (!_current_block) throw std::runtime_error("No data block to release");
    _parent._ring->commit_read(false, _index);
    _current_block.reset();
    ++_block_idx;
}

bool DadaReadClient::DataStream::at_end() const { return _parent._ring->at_end(false); }

}  // namespace psrdada_cpp
