// Standard MIDI File loader for render jobs.
// Behaviour follows src/io/midi.h:8-26 and src/io/midi.cpp of the reference.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

namespace clay {

enum class TailPolicy { natural, orig, custom };

enum class EventKind { on, off, control, bend };

// One channel event at an output frame (src/io/midi.h:12-17).
struct Event {
    int frame;
    EventKind kind;
    int a;  // key, controller number, or signed 14-bit bend value
    int b;  // velocity or controller value
};

// Frame-ordered events, the end marker frame, and the optional QA measures
// (src/io/midi.h:19-24).
struct Midi {
    std::vector<Event> events;
    int end = 0;
    int bodyFrames = 0;
    std::vector<int> onsets;
};

// Parse SMF bytes (src/io/midi.cpp:246-327). Throws std::runtime_error.
Midi parseMidi(const std::string& bytes, int rate, TailPolicy tail, bool measure);

// Read and parse a file; errors carry the path (src/io/midi.cpp:331-342).
Midi loadMidi(const std::string& path, int rate, TailPolicy tail, bool measure = false);

} // namespace clay
