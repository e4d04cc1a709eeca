// ibp_io.hip -- host file loader behind read_shared.
//
// Replaces the copy loop of read_shared (src/ibp_api.cu:52-122): the reference
// mmaps the file in 20 GiB pieces and memcpys each into a shared-memory
// tensor, one thread, and copies st_size bytes whatever the tensor's size.
// Here the file is read with pread by several threads in 64 MiB pieces straight
// into the caller's buffer (the tensor's shared-memory storage), and a file
// larger than the buffer is an error instead of an overflow.
#include <errno.h>
#include <fcntl.h>
#include <string.h>
#include <sys/stat.h>
#include <unistd.h>

#include <atomic>
#include <thread>
#include <vector>

#include "ibp_internal.h"

using namespace ibp;

extern "C" int ibp_read_file(const char *path, void *dst, size_t capacity, int threads, size_t *bytes_read)
{
    IBP_TRACE("ibp_read_file");
    IBP_CHECK(path != nullptr, "null path");
    if (bytes_read) *bytes_read = 0;
    const int fd = open(path, O_RDONLY | O_CLOEXEC);
    if (fd < 0) return set_error(IBP_ERR_INVALID, "cannot open %s: %s", path, strerror(errno));
    struct stat sb;
    if (fstat(fd, &sb) != 0) {
        const int e = errno;
        close(fd);
        return set_error(IBP_ERR_INVALID, "cannot stat %s: %s", path, strerror(e));
    }
    const size_t size = (size_t)sb.st_size;
    if (size > capacity) {
        close(fd);
        return set_error(IBP_ERR_INVALID, "%s holds %zu bytes, more than the %zu bytes of the tensor", path, size,
                         capacity);
    }
    if (size > 0 && dst == nullptr) {
        close(fd);
        return set_error(IBP_ERR_INVALID, "null destination");
    }
    (void)posix_fadvise(fd, 0, 0, POSIX_FADV_SEQUENTIAL);
    constexpr size_t kPiece = 64ull << 20;
    const size_t pieces = (size + kPiece - 1) / kPiece;
    int nthreads = threads > 0 ? (threads < 64 ? threads : 64) : 16;
    if ((size_t)nthreads > pieces) nthreads = pieces > 0 ? (int)pieces : 1;

    std::atomic<size_t> next{0};
    std::atomic<int> failed{0};
    auto worker = [&]() {
        for (;;) {
            const size_t k = next.fetch_add(1);
            if (k >= pieces || failed.load()) return;
            size_t off = k * kPiece;
            const size_t end = off + kPiece < size ? off + kPiece : size;
            while (off < end) {
                const ssize_t got = pread(fd, (char *)dst + off, end - off, (off_t)off);
                if (got < 0 && errno == EINTR) continue;
                if (got <= 0) { // error, or the file shrank while being read
                    failed.store(got < 0 ? errno : EIO);
                    return;
                }
                off += (size_t)got;
            }
        }
    };
    std::vector<std::thread> pool;
    for (int t = 1; t < nthreads; ++t) pool.emplace_back(worker);
    worker();
    for (auto &t : pool) t.join();
    close(fd);
    if (failed.load()) return set_error(IBP_ERR_INVALID, "reading %s failed: %s", path, strerror(failed.load()));
    if (bytes_read) *bytes_read = size;
    return IBP_OK;
}
