#include "audio.h"

#include <algorithm>
#include <atomic>
#include <cerrno>
#include <cstring>
#include <fcntl.h>
#include <filesystem>
#include <fstream>
#include <iterator>
#include <limits>
#include <stdexcept>
#include <unistd.h>

namespace clay {
namespace fs = std::filesystem;

namespace {

uint32_t le32(const unsigned char* p) { return p[0] | (p[1] << 8) | (p[2] << 16) | (uint32_t(p[3]) << 24); }
uint16_t le16(const unsigned char* p) { return uint16_t(p[0] | (p[1] << 8)); }

std::string readAll(const std::string& path)
{
    std::ifstream in(path, std::ios::binary);
    if (!in)
        return {};
    return std::string{std::istreambuf_iterator<char>(in), {}};
}

} // namespace

WaveData readWave(const std::string& path, bool withData)
{
    const auto bytes = readAll(path);
    const auto* p = reinterpret_cast<const unsigned char*>(bytes.data());
    const size_t size = bytes.size();
    const auto bad = [&]() -> WaveData { throw std::runtime_error("cannot decode mono or stereo sample " + path); };
    if (size < 12 || std::memcmp(p, "RIFF", 4) || std::memcmp(p + 8, "WAVE", 4))
        return bad();
    WaveData wave;
    int bits = 0, format = 0;
    const unsigned char* data = nullptr;
    size_t dataBytes = 0;
    for (size_t at = 12; at + 8 <= size;) {
        const uint32_t length = le32(p + at + 4);
        const auto* body = p + at + 8;
        const size_t avail = std::min<size_t>(length, size - at - 8);
        if (!std::memcmp(p + at, "fmt ", 4) && avail >= 16) {
            format = le16(body);
            wave.channels = le16(body + 2);
            wave.rate = int(le32(body + 4));
            bits = le16(body + 14);
            if (format == 0xfffe && avail >= 26)
                format = le16(body + 24);  // WAVE_FORMAT_EXTENSIBLE sub-format
        } else if (!std::memcmp(p + at, "data", 4)) {
            data = body;
            dataBytes = avail;
        } else if (!std::memcmp(p + at, "smpl", 4) && avail >= 36) {
            const uint32_t loops = le32(body + 28);
            if (loops > 0 && avail >= 36 + 24) {
                const auto* loop = body + 36;
                const uint32_t type = le32(loop + 4), start = le32(loop + 8), end = le32(loop + 12);
                if (type == 0 && start <= end) {  // forward loop; the chunk stores an inclusive end
                    wave.loopFirst = int(start);
                    wave.loopEnd = int(end) + 1;
                }
            }
        }
        at += 8 + length + (length & 1);
    }
    if (format != 1 || bits != 16 || (wave.channels != 1 && wave.channels != 2) || wave.rate <= 0 || !data)
        return bad();
    const size_t frames = dataBytes / (2 * size_t(wave.channels));
    if (frames == 0 || frames > size_t(std::numeric_limits<int>::max()))
        return bad();
    wave.frames = int(frames);
    if (!(wave.loopFirst < wave.loopEnd && wave.loopEnd <= wave.frames))
        wave.loopFirst = wave.loopEnd = 0;
    if (withData) {
        wave.pcm.resize(frames * wave.channels);
        for (size_t i = 0; i < wave.pcm.size(); ++i)
            wave.pcm[i] = int16_t(le16(data + 2 * i));
    }
    return wave;
}

int AudioBank::describe(const std::string& path)
{
    if (const auto it = loaded_.find(path); it != loaded_.end())
        return it->second;
    const auto key = fs::weakly_canonical(fs::path(path)).string();
    if (const auto it = loaded_.find(key); it != loaded_.end()) {
        loaded_.emplace(path, it->second);
        return it->second;
    }
    const auto wave = readWave(key, false);
    const int id = int(samples.size());
    samples.push_back({UINT64_MAX, wave.frames, wave.rate, wave.channels, wave.loopFirst, wave.loopEnd});
    paths_.push_back(key);
    loaded_.emplace(key, id);
    if (path != key)
        loaded_.emplace(path, id);
    return id;
}

const Sample& AudioBank::decode(int id)
{
    auto& sample = samples.at(size_t(id));
    if (sample.offset != UINT64_MAX)
        return sample;
    const auto& path = paths_.at(size_t(id));
    const auto wave = readWave(path, true);
    if (wave.frames != sample.frames || wave.rate != sample.rate || wave.channels != sample.channels)
        throw std::runtime_error("sample changed since it was opened " + path);
    sample.offset = pcm.size();
    pcm.insert(pcm.end(), wave.pcm.begin(), wave.pcm.end());
    return sample;
}

void AudioBank::releaseDecoded()
{
    std::vector<int16_t>().swap(pcm);
    for (auto& sample : samples)
        sample.offset = UINT64_MAX;
}

std::vector<unsigned char> waveHeader(uint32_t dataBytes, int rate, int channels)
{
    std::vector<unsigned char> h(44, 0);
    const auto put = [&](size_t at, uint32_t v, int n) {
        for (int i = 0; i < n; ++i)
            h[at + i] = static_cast<unsigned char>(v >> (8 * i));
    };
    std::memcpy(h.data(), "RIFF", 4);
    put(4, dataBytes + 36, 4);
    std::memcpy(h.data() + 8, "WAVEfmt ", 8);
    put(16, 16, 4);
    put(20, 1, 2);
    put(22, uint32_t(channels), 2);
    put(24, uint32_t(rate), 4);
    put(28, uint32_t(rate * channels * 2), 4);
    put(32, uint32_t(channels * 2), 2);
    put(34, 16, 2);
    std::memcpy(h.data() + 36, "data", 4);
    put(40, dataBytes, 4);
    return h;
}

void writeWave(const std::string& path, const int16_t* pcm, int dataFrames, int frames, int rate, int channels)
{
    const uint64_t stride = uint64_t(channels) * 2;
    if (frames <= 0 || dataFrames < 0 || dataFrames > frames || uint64_t(frames) * stride + 36 > UINT32_MAX)
        throw std::runtime_error("output exceeds PCM16 WAV limits");
    std::ofstream out(path, std::ios::binary);
    const auto header = waveHeader(uint32_t(uint64_t(frames) * stride), rate, channels);
    out.write(reinterpret_cast<const char*>(header.data()), std::streamsize(header.size()));
    if (dataFrames)
        out.write(reinterpret_cast<const char*>(pcm), std::streamsize(uint64_t(dataFrames) * stride));
    static const char zeros[16384] = {};
    for (uint64_t left = uint64_t(frames - dataFrames) * stride; left;) {
        const auto n = std::min<uint64_t>(left, sizeof zeros);
        out.write(zeros, std::streamsize(n));
        left -= n;
    }
    out.close();
    if (!out)
        throw std::runtime_error("cannot write WAV " + path);
}

bool completeOutput(const std::string& path, bool flac, int rate, int channels)
{
    if (!flac) {
        std::ifstream in(path, std::ios::binary | std::ios::ate);
        if (!in)
            return false;
        const std::streamoff size = in.tellg();
        if (size <= 44 || size - 44 > std::streamoff(UINT32_MAX) - 36 || (size - 44) % (channels * 2))
            return false;
        std::vector<unsigned char> header(44);
        in.seekg(0);
        return bool(in.read(reinterpret_cast<char*>(header.data()), 44)) &&
               header == waveHeader(uint32_t(size - 44), rate, channels);
    }
    int fileRate = 0, fileChannels = 0;
    std::vector<int16_t> pcm;
    return readFlac(path, fileRate, fileChannels, pcm) && fileRate == rate && fileChannels == channels && !pcm.empty();
}

PendingOutput::PendingOutput(const std::string& destination) : destination_(destination)
{
    static std::atomic<uint64_t> serial{0};
    const auto dir = fs::path(destination_).parent_path();
    for (;;) {
        temporary_ = (dir / ("claysampler." + std::to_string(getpid()) + "." +
                             std::to_string(serial.fetch_add(1)) + ".part")).string();
        const int fd = ::open(temporary_.c_str(), O_WRONLY | O_CREAT | O_EXCL, 0666);
        if (fd >= 0) {
            ::close(fd);
            break;
        }
        if (errno != EEXIST)
            throw std::runtime_error("cannot create temporary output for " + destination_);
    }
}

PendingOutput::~PendingOutput()
{
    std::error_code ec;
    if (!temporary_.empty())
        fs::remove(temporary_, ec);
}

void PendingOutput::publish()
{
    const auto sync = [](const std::string& path, int flags) {
        const int fd = ::open(path.c_str(), flags);
        if (fd < 0)
            throw std::runtime_error("cannot open for sync " + path);
        const int status = ::fsync(fd);
        const int closed = ::close(fd);
        if (status || closed)
            throw std::runtime_error("cannot sync " + path);
    };
    sync(temporary_, O_WRONLY);
    fs::rename(temporary_, destination_);
    const auto parent = fs::path(destination_).parent_path();
    sync(parent.empty() ? std::string(".") : parent.string(), O_RDONLY | O_DIRECTORY);
    temporary_.clear();
}

} // namespace clay
