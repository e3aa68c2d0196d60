// Standard MIDI File loader. The observable behaviour (accepted files, event
// frames, natural/custom preparation, QA onsets) follows src/io/midi.cpp of
// the reference; the code is this project's own.
#include "midi.h"

#include <algorithm>
#include <array>
#include <climits>
#include <cmath>
#include <fstream>
#include <iterator>
#include <stdexcept>

namespace clay {
namespace {

[[noreturn]] void fail(const char* message) { throw std::runtime_error(message); }

// Byte cursor over an SMF chunk.
struct Cursor {
    const unsigned char* at;
    const unsigned char* stop;

    size_t left() const { return size_t(stop - at); }
    unsigned u8()
    {
        if (at == stop)
            fail("truncated MIDI");
        return *at++;
    }
    uint32_t be(int bytes)
    {
        uint32_t v = 0;
        for (int i = 0; i < bytes; ++i)
            v = (v << 8) | u8();
        return v;
    }
    uint32_t varlen()  // at most four bytes (src/io/midi.cpp:57-67)
    {
        uint32_t v = 0;
        for (int i = 0; i < 4; ++i) {
            const unsigned byte = u8();
            v = (v << 7) | (byte & 0x7f);
            if (byte < 0x80)
                return v;
        }
        fail("invalid MIDI variable-length value");
    }
    Cursor sub(size_t bytes)
    {
        if (bytes > left())
            fail("truncated MIDI");
        Cursor c{at, at + bytes};
        at += bytes;
        return c;
    }
};

// A raw message. status 0xff = tempo (a = microseconds per quarter),
// status 0 = end-of-track marker, otherwise a channel message.
struct Raw {
    uint64_t tick;
    int status, a, b;
};

int kindOf(const Raw& m) { return m.status >> 4; }
bool isNoteOn(const Raw& m) { return kindOf(m) == 9 && m.b != 0; }
bool isNote(const Raw& m) { return kindOf(m) == 8 || kindOf(m) == 9; }

// src/io/midi.cpp:89-132
void readTrack(Cursor track, std::vector<Raw>& out)
{
    uint64_t tick = 0;
    int running = 0;
    while (track.left()) {
        tick += track.varlen();
        int status = int(track.u8());
        int first = -1;
        if (status < 0x80) {
            first = status;
            status = running;
            if (!status)
                fail("MIDI running status has no channel message");
        }
        if (status == 0xff) {
            const unsigned type = track.u8();
            Cursor body = track.sub(track.varlen());
            if (type == 0x51) {
                const uint32_t tempo = body.be(3);
                if (!tempo || body.left())
                    fail("invalid MIDI tempo");
                out.push_back({tick, 0xff, int(tempo), 0});
            } else if (type == 0x2f) {
                if (body.left())
                    fail("invalid MIDI end-of-track");
                out.push_back({tick, 0, 0, 0});
                return;
            }
            running = 0;
            continue;
        }
        if (status >= 0xf0)
            fail("unsupported MIDI system or SysEx message");
        running = status;
        const int a = first >= 0 ? first : int(track.u8());
        const int kind = status >> 4;
        const int b = (kind == 0xc || kind == 0xd) ? 0 : int(track.u8());
        if (a > 0x7f || b > 0x7f)
            fail("invalid MIDI data byte");
        out.push_back({tick, status, a, b});
    }
    fail("missing MIDI end-of-track");
}

// Sort rank at equal ticks after natural preparation: controllers and tempo,
// then note-offs, then note-ons (src/io/midi.cpp:134-141).
int rank(const Raw& m)
{
    if (isNoteOn(m))
        return 30;
    return isNote(m) ? 25 : 15;
}

// Natural policy: end at the last completed note or the release of a pedal
// held across it, drop trailing events and panic commands, add the
// ClayEngine controller defaults (src/io/midi.cpp:143-203).
void prepareNatural(std::vector<Raw>& messages)
{
    std::vector<int> held(16 * 128, 0);
    uint64_t lastNote = 0;
    int channel = -1;
    for (const auto& m : messages) {
        if (!isNote(m))
            continue;
        const int ch = m.status & 15;
        int& count = held[ch * 128 + m.a];
        if (isNoteOn(m)) {
            ++count;
        } else if (count > 0) {
            --count;
            if (channel >= 0 && channel != ch)
                fail("multiple MIDI channels in one render job");
            channel = ch;
            lastNote = m.tick;
        }
    }
    if (channel < 0)
        fail("natural rendering requires a complete MIDI note");

    static const int pedals[3] = {64, 66, 69};
    bool down[3] = {false, false, false};
    uint64_t end = lastNote;
    for (const auto& m : messages) {
        if (kindOf(m) != 0xb)
            continue;
        for (int p = 0; p < 3; ++p) {
            if (m.a != pedals[p])
                continue;
            if (m.tick <= end) {
                down[p] = m.b >= 64;
            } else if (down[p] && m.b < 64) {
                end = m.tick;
                down[p] = false;
            }
        }
    }
    const auto dropped = [&](const Raw& m) {
        if (m.status == 0 || m.tick > end || kindOf(m) == 0xc)
            return true;
        if (kindOf(m) != 0xb)
            return false;
        switch (m.a) {
        case 0: case 7: case 10: case 32: case 91: case 93:
            return true;
        case 120: case 123:
            return m.tick >= lastNote;
        default:
            return false;
        }
    };
    messages.erase(std::remove_if(messages.begin(), messages.end(), dropped), messages.end());
    const int cc = 0xb0 | channel;
    messages.insert(messages.begin(), {Raw{0, cc, 7, 100}, Raw{0, cc, 10, 64}});
    std::stable_sort(messages.begin(), messages.end(), [](const Raw& x, const Raw& y) {
        return x.tick != y.tick ? x.tick < y.tick : rank(x) < rank(y);
    });
    for (int pedal : pedals)
        messages.push_back({end, cc, pedal, 0});
    messages.push_back({end, 0, 0, 0});
}

// QA inputs: unique note-on frames of completed notes and the last note-off
// frame (src/io/midi.cpp:205-244). Runs on the time-sorted raw messages.
void measureNotes(const std::vector<Raw>& messages, int rate, uint32_t division, Midi& midi)
{
    // FIFO of pending note-on frames per (channel, key).
    std::vector<std::vector<int>> open(16 * 128);
    std::vector<size_t> head(16 * 128, 0);
    int tempo = 500000;
    uint64_t previous = 0;
    double seconds = 0;
    for (const auto& m : messages) {
        seconds += double(m.tick - previous) * (tempo * 1e-6 / division);
        previous = m.tick;
        if (m.status == 0xff) {
            tempo = m.a;
            continue;
        }
        if (!isNote(m))
            continue;
        if (seconds * rate >= INT_MAX / 2)
            fail("MIDI exceeds maximum render duration");
        const int frame = int(std::nearbyint(seconds * rate));
        const int slot = (m.status & 15) * 128 + m.a;
        if (isNoteOn(m)) {
            open[slot].push_back(frame);
        } else if (head[slot] < open[slot].size()) {
            midi.onsets.push_back(open[slot][head[slot]++]);
            midi.bodyFrames = std::max(midi.bodyFrames, frame);
        }
    }
    std::sort(midi.onsets.begin(), midi.onsets.end());
    midi.onsets.erase(std::unique(midi.onsets.begin(), midi.onsets.end()), midi.onsets.end());
}

} // namespace

Midi parseMidi(const std::string& bytes, int rate, TailPolicy tail, bool measure)
{
    const auto* base = reinterpret_cast<const unsigned char*>(bytes.data());
    Cursor file{base, base + bytes.size()};
    const auto tag = [&](const char* expected) {
        Cursor id = file.sub(4);
        return std::equal(id.at, id.stop, expected);
    };
    if (!tag("MThd"))
        fail("expected a Standard MIDI File");
    Cursor header = file.sub(file.be(4));
    const uint32_t format = header.be(2), tracks = header.be(2), division = header.be(2);
    if (format > 1 || tracks == 0 || (format == 0 && tracks != 1) || division == 0 || division >= 32768)
        fail("expected format 0/1 MIDI with PPQ timing");
    std::vector<Raw> messages;
    for (uint32_t t = 0; t < tracks; ++t) {
        if (!tag("MTrk"))
            fail("missing MIDI track chunk");
        readTrack(file.sub(file.be(4)), messages);
    }
    std::stable_sort(messages.begin(), messages.end(), [](const Raw& x, const Raw& y) { return x.tick < y.tick; });

    Midi midi;
    if (measure)
        measureNotes(messages, rate, division, midi);
    if (tail == TailPolicy::natural)
        prepareNatural(messages);
    if (tail == TailPolicy::custom) {
        // Cut after the last performance message; release the sustain pedal there
        // (src/io/midi.cpp:270-282).
        auto last = std::find_if(messages.rbegin(), messages.rend(), [](const Raw& m) {
            const int k = kindOf(m);
            return k == 8 || k == 9 || k == 0xb || k == 0xe;
        });
        if (last != messages.rend()) {
            const Raw kept = *last;
            messages.erase(last.base(), messages.end());
            messages.push_back({kept.tick, 0xb0 | (kept.status & 15), 64, 0});
            messages.push_back({kept.tick, 0, 0, 0});
        }
    }

    // Tick -> frame with tempo changes (src/io/midi.cpp:283-326).
    const double scale = rate / (1000000.0 * division);
    double frames = 0;
    uint64_t previous = 0;
    int tempo = 500000, channel = -1;
    for (const auto& m : messages) {
        frames += double(m.tick - previous) * tempo * scale;
        previous = m.tick;
        if (frames >= INT_MAX / 2)
            fail("MIDI exceeds maximum render duration");
        const int frame = int(tail == TailPolicy::natural ? std::nearbyint(frames) : std::floor(frames + 1e-7));
        if (m.status == 0xff) {
            tempo = m.a;
            continue;
        }
        if (m.status == 0) {
            midi.end = frame;
            continue;
        }
        const int kind = kindOf(m), ch = m.status & 15;
        if (channel < 0)
            channel = ch;
        if (channel != ch)
            fail("multiple MIDI channels in one render job");
        if (kind == 8 || kind == 9)
            midi.events.push_back({frame, isNoteOn(m) ? EventKind::on : EventKind::off, m.a, m.b});
        else if (kind == 0xb)
            midi.events.push_back({frame, EventKind::control, m.a, m.b});
        else if (kind == 0xe)
            midi.events.push_back({frame, EventKind::bend, (m.a | (m.b << 7)) - 8192, 0});
        else if (kind != 0xc)  // program change: the job selects the preset
            throw std::runtime_error("unsupported MIDI channel message type " + std::to_string(kind));
    }
    return midi;
}

Midi loadMidi(const std::string& path, int rate, TailPolicy tail, bool measure)
{
    std::ifstream in(path, std::ios::binary);
    if (!in)
        throw std::runtime_error("cannot open MIDI " + path);
    const std::string bytes{std::istreambuf_iterator<char>(in), {}};
    try {
        return parseMidi(bytes, rate, tail, measure);
    } catch (const std::exception& e) {
        throw std::runtime_error(path + ": " + e.what());
    }
}

} // namespace clay
