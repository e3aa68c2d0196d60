// Test infrastructure: a C entry point over the reference's own MIDI loader
// (src/io/midi.cpp, compiled from its source file by oracle/Makefile into
// oracle/_ref/). Used only by tests to pin this project's MIDI loader.
#include "midi.h"

#include <cstring>
#include <exception>
#include <string>

extern "C" int clay_ref_load_midi(const char* path, int rate, int tail, int measure, int* events, int capacity,
                                  int* count, int* end, int* bodyFrames, int* onsets, int onsetCapacity,
                                  int* onsetCount, char* error, int errorCapacity)
{
    try {
        const auto midi = clay::loadMidi(path, rate, static_cast<clay::TailPolicy>(tail), measure != 0);
        *count = int(midi.events.size());
        *end = midi.end;
        *bodyFrames = midi.bodyFrames;
        *onsetCount = int(midi.onsets.size());
        for (int i = 0; i < *count && i < capacity; ++i) {
            const auto& e = midi.events[size_t(i)];
            events[4 * i] = e.frame;
            events[4 * i + 1] = int(e.kind);
            events[4 * i + 2] = e.a;
            events[4 * i + 3] = e.b;
        }
        for (int i = 0; i < *onsetCount && i < onsetCapacity; ++i)
            onsets[i] = midi.onsets[size_t(i)];
        return 0;
    } catch (const std::exception& e) {
        std::strncpy(error, e.what(), size_t(errorCapacity - 1));
        error[errorCapacity - 1] = 0;
        return -1;
    }
}
