// Writer of a DADA ring (dada_write_client.hpp:10-70).
#ifndef PSRDADA_CPP_DADA_WRITE_CLIENT_HPP
#define PSRDADA_CPP_DADA_WRITE_CLIENT_HPP

#include <memory>

#include "psrdada_cpp/dada_ring.hpp"
#include "psrdada_cpp/multilog.hpp"
#include "psrdada_cpp/raw_bytes.hpp"

namespace psrdada_cpp {

class DadaWriteClient {
public:
    class HeaderStream {
    public:
        explicit HeaderStream(DadaWriteClient& parent) : _parent(parent) {}
        HeaderStream(HeaderStream const&) = delete;
        RawBytes& next();
        void release();

    private:
        DadaWriteClient& _parent;
        std::unique_ptr<RawBytes> _current_block;
        std::size_t _index = 0;
    };

    class DataStream {
    public:
        explicit DataStream(DadaWriteClient& parent) : _parent(parent) {}
        DataStream(DataStream const&) = delete;
        RawBytes& next();
        // eod = true marks this block as the last of the observation.
        void release(bool eod = false);
        std::size_t block_idx() const { return _block_idx; }

    private:
        DadaWriteClient& _parent;
        std::unique_ptr<RawBytes> _current_block;
        std::size_t _index = 0;
        std::size_t _block_idx = 0;
    };

    DadaWriteClient(key_t key, MultiLog& log);
    DadaWriteClient(DadaWriteClient const&) = delete;
    ~DadaWriteClient();

    HeaderStream& header_stream() { return _header_stream; }
    DataStream& data_stream() { return _data_stream; }
    std::size_t data_buffer_size() const { return _ring->data_buffer_size(); }
    std::size_t header_buffer_size() const { return _ring->header_buffer_size(); }
    std::size_t data_buffer_count() const { return _ring->data_buffer_count(); }
    void reset();

private:
    friend class HeaderStream;
    friend class DataStream;
    key_t _key;
    MultiLog& _log;
    std::shared_ptr<DadaRing> _ring;
    HeaderStream _header_stream;
    DataStream _data_stream;
};

}  // namespace psrdada_cpp

#endif
