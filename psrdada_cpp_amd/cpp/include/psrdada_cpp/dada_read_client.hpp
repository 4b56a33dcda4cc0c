// Reader of a DADA ring (dada_read_client.hpp:14-160).
#ifndef PSRDADA_CPP_DADA_READ_CLIENT_HPP
#define PSRDADA_CPP_DADA_READ_CLIENT_HPP

#include <memory>

#include "psrdada_cpp/dada_ring.hpp"
#include "psrdada_cpp/multilog.hpp"
#include "psrdada_cpp/raw_bytes.hpp"

namespace psrdada_cpp {

class DadaReadClient {
public:
    class HeaderStream {
    public:
        explicit HeaderStream(DadaReadClient& parent) : _parent(parent) {}
        HeaderStream(HeaderStream const&) = delete;
        RawBytes& next();
        void release();
        bool at_end() const;

    private:
        DadaReadClient& _parent;
        std::unique_ptr<RawBytes> _current_block;
        std::size_t _index = 0;
    };

    class DataStream {
    public:
        explicit DataStream(DadaReadClient& parent) : _parent(parent) {}
        DataStream(DataStream const&) = delete;
        RawBytes& next();
        void release();
        bool at_end() const;
        std::size_t block_idx() const { return _block_idx; }

    private:
        DadaReadClient& _parent;
        std::unique_ptr<RawBytes> _current_block;
        std::size_t _index = 0;
        std::size_t _block_idx = 0;
    };

    DadaReadClient(key_t key, MultiLog& log);
    DadaReadClient(DadaReadClient const&) = delete;
    ~DadaReadClient();

    HeaderStream& header_stream() { return _header_stream; }
    DataStream& data_stream() { return _data_stream; }
    std::size_t data_buffer_size() const { return _ring->data_buffer_size(); }
    std::size_t header_buffer_size() const { return _ring->header_buffer_size(); }

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
