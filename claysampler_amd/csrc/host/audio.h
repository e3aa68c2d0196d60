// Sample bank and PCM16 output files.
// Interface after src/io/audio.h of the reference; the libsndfile decoder it
// loads at run time is replaced by this project's own PCM16 WAV reader and the
// WAV/FLAC writers are this project's own.
#pragma once

#include <cstdint>
#include <map>
#include <string>
#include <vector>

namespace clay {

// Decoded sample location in the bank (src/io/audio.h:9-16).
struct Sample {
    uint64_t offset;  // UINT64_MAX until decoded
    int frames, rate, channels, loopFirst, loopEnd;
};

// Append-only int16 sample bank (src/io/audio.h:18-37). int16 v here equals the
// reference's float v/32768 exactly.
class AudioBank {
public:
    int describe(const std::string& path);
    const Sample& decode(int id);
    void releaseDecoded();

    std::vector<Sample> samples;
    std::vector<int16_t> pcm;

private:
    std::map<std::string, int> loaded_;
    std::vector<std::string> paths_;
};

// PCM16 WAV contents.
struct WaveData {
    int rate = 0, channels = 0, frames = 0;
    int loopFirst = 0, loopEnd = 0;  // first forward loop of a `smpl` chunk, end exclusive
    std::vector<int16_t> pcm;
};
WaveData readWave(const std::string& path, bool withData);

// 44-byte canonical header (src/io/audio.cpp:38-58).
std::vector<unsigned char> waveHeader(uint32_t dataBytes, int rate, int channels);

// Write `dataFrames` frames then zero frames up to `frames` (src/io/audio.cpp:264-285).
void writeWave(const std::string& path, const int16_t* pcm, int dataFrames, int frames, int rate, int channels);

// FLAC (RFC 9639) with fixed and LPC-free verbatim/fixed subframes, Rice-coded
// residuals; `level` 0..8 selects the search effort. Replaces libsndfile's
// encoder (src/io/audio.cpp:236-262).
void writeFlac(const std::string& path, const int16_t* pcm, int dataFrames, int frames, int rate, int channels,
               int level);

// Decode a FLAC written by any encoder that uses the subset features this
// decoder supports (fixed, verbatim, constant and LPC subframes; 16-bit).
bool readFlac(const std::string& path, int& rate, int& channels, std::vector<int16_t>& pcm);

// --resume completeness check (src/io/audio.cpp:156-180).
bool completeOutput(const std::string& path, bool flac, int rate, int channels);

// Temporary sibling file, published by fsync + rename + directory fsync
// (src/io/audio.cpp:182-234).
class PendingOutput {
public:
    explicit PendingOutput(const std::string& destination);
    ~PendingOutput();
    PendingOutput(const PendingOutput&) = delete;
    PendingOutput& operator=(const PendingOutput&) = delete;
    const std::string& path() const { return temporary_; }
    void publish();

private:
    std::string destination_, temporary_;
};

} // namespace clay
