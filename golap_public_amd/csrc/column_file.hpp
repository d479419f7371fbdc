// Column file layout shared by the loader (storage.cpp) and the streaming scans
// (outofcore.cpp): a header, the catalog and zone maps, then the page data as one
// 4 KiB-aligned region padded to a 4 KiB boundary.
#pragma once

#include <atomic>
#include <cstdint>
#include <string>
#include <vector>

#include "host_column.hpp"

namespace golap_file {

constexpr uint64_t ALIGN = 4096;

struct FileHeader {
    char magic[8];
    uint32_t version, type, elem, page_rows;
    uint64_t chunk_rows, n_rows, n_pages, n_chunks, comp_bytes, payload_bytes;
    uint64_t data_offset;  // 4 KiB-aligned start of the page data
    uint64_t file_bytes;   // data_offset + comp_bytes rounded up to 4 KiB
};
static_assert(sizeof(FileHeader) == 88, "file header layout");

uint64_t align_up(uint64_t v);
uint64_t catalog_bytes(uint64_t n_pages, uint64_t n_chunks);
bool write_all(int fd, const void *p, uint64_t n, uint64_t off);
// pread until n bytes are read; false on an error or end of file
bool read_all(int fd, void *p, uint64_t n, uint64_t off);
// Header and catalog of a column file into h (page data left empty); false with err set.
bool read_catalog(int fd, golap_host_column &h, FileHeader &fh, std::string &err);

// A read of len bytes at file offset at into dst: through dfd (O_DIRECT; whole 4 KiB blocks into
// a 4 KiB-aligned dst) while *direct holds, else through bfd (buffered). A file system that
// accepts O_DIRECT at open but refuses the reads (stacked file systems) clears *direct and the
// read is retried buffered.
struct FileRead {
    int dfd, bfd;
    std::atomic<bool> *direct;
    uint64_t at, len;
    uint8_t *dst;
};
constexpr uint64_t READ_PIECE = 8ull << 20;
constexpr unsigned READ_THREADS = 8;
// Appends reads of [at, at + len) cut into READ_PIECE pieces.
void split_read(std::vector<FileRead> &out, int dfd, int bfd, std::atomic<bool> *direct, uint64_t at, uint64_t len,
                uint8_t *dst);
// Runs every read with up to READ_THREADS threads; false if one failed.
bool read_parallel(const std::vector<FileRead> &reads);

}  // namespace golap_file
