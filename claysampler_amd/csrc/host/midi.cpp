// Standard MIDI File reader for render jobs.
//
// The file is decoded in three stages, each a function of its own:
//   1. decodeTracks: every track's messages, merged into one tick-ordered list
//      (stable across tracks, as src/io/midi.cpp:235-236 orders them);
//   2. the tail policy's edit of that list: naturalEnding (src/io/midi.cpp:
//      115-175) or customEnding (:247-260);
//   3. FrameClock: ticks to output frames through the tempo changes
//      (:262-299), accumulated message by message so that every frame rounds
//      exactly as the reference's running sum does.
// The QA measures (note onsets, body end; :177-216) read the list between
// stages 1 and 2. The accepted files, error texts, event frames and their
// order are the reference's interface and are kept; the code is this
// project's own.
#include "midi.h"

#include <algorithm>
#include <climits>
#include <cmath>
#include <deque>
#include <fstream>
#include <iterator>
#include <stdexcept>

namespace clay {
namespace {

[[noreturn]] void reject(const std::string& why) { throw std::runtime_error(why); }

// A bounded stretch of the file.
struct Bytes {
    const unsigned char* at;
    const unsigned char* end;
    size_t size() const { return size_t(end - at); }
};

unsigned nextByte(Bytes& in)
{
    if (in.at == in.end)
        reject("truncated MIDI");
    return *in.at++;
}

uint32_t bigEndian(Bytes& in, int count)
{
    uint32_t v = 0;
    while (count-- > 0)
        v = v << 8 | nextByte(in);
    return v;
}

// Variable-length quantity of at most four bytes.
uint32_t quantity(Bytes& in)
{
    uint32_t v = 0;
    for (int used = 1;; ++used) {
        const unsigned byte = nextByte(in);
        v = v << 7 | (byte & 0x7f);
        if (!(byte & 0x80))
            return v;
        if (used == 4)
            reject("invalid MIDI variable-length value");
    }
}

Bytes chunk(Bytes& in, size_t length)
{
    if (length > in.size())
        reject("truncated MIDI");
    Bytes part{in.at, in.at + length};
    in.at += length;
    return part;
}

bool chunkIs(Bytes& in, const char (&tag)[5])
{
    const Bytes id = chunk(in, 4);
    return std::equal(id.at, id.end, tag);
}

// One message of any track at its absolute tick.
struct Message {
    enum Kind { channel, tempo, trackEnd } kind;
    uint64_t tick;
    int status;         // channel messages: status byte
    int first, second;  // data bytes (second 0 for one-byte messages)
    int microseconds;   // tempo: microseconds per quarter note

    int type() const { return kind == channel ? status >> 4 : -1; }
    int channelIndex() const { return status & 15; }
    bool sounding() const { return type() == 9 && second != 0; }
    bool noteEnd() const { return (type() == 8 || type() == 9) && !sounding(); }
    bool note() const { return type() == 8 || type() == 9; }
};

Message channelMessage(uint64_t tick, int status, int first, int second)
{
    return {Message::channel, tick, status, first, second, 0};
}

// Messages of one MTrk chunk (src/io/midi.cpp:61-104). Meta events other than
// tempo and end-of-track are skipped; they and SysEx clear running status.
void decodeTrack(Bytes track, std::vector<Message>& out)
{
    uint64_t tick = 0;
    int running = 0;
    while (track.size()) {
        tick += quantity(track);
        const unsigned lead = nextByte(track);
        if (lead == 0xff) {
            const unsigned meta = nextByte(track);
            Bytes body = chunk(track, quantity(track));
            running = 0;
            if (meta == 0x2f) {
                if (body.size())
                    reject("invalid MIDI end-of-track");
                out.push_back({Message::trackEnd, tick, 0, 0, 0, 0});
                return;  // bytes after end-of-track are not read
            }
            if (meta == 0x51) {
                const uint32_t quarter = bigEndian(body, 3);
                if (quarter == 0 || body.size())
                    reject("invalid MIDI tempo");
                out.push_back({Message::tempo, tick, 0, 0, 0, int(quarter)});
            }
            continue;
        }
        if (lead >= 0xf0)
            reject("unsupported MIDI system or SysEx message");
        int status = int(lead), first;
        if (lead & 0x80) {
            first = int(nextByte(track));
        } else {
            if (!running)
                reject("MIDI running status has no channel message");
            status = running;
            first = int(lead);
        }
        running = status;
        const int type = status >> 4;
        const int second = type == 0xc || type == 0xd ? 0 : int(nextByte(track));
        if (first > 127 || second > 127)
            reject("invalid MIDI data byte");
        out.push_back(channelMessage(tick, status, first, second));
    }
    reject("missing MIDI end-of-track");
}

std::vector<Message> decodeTracks(const std::string& data, int& division)
{
    Bytes file{reinterpret_cast<const unsigned char*>(data.data()),
               reinterpret_cast<const unsigned char*>(data.data()) + data.size()};
    if (!chunkIs(file, "MThd"))
        reject("expected a Standard MIDI File");
    Bytes header = chunk(file, bigEndian(file, 4));
    const uint32_t format = bigEndian(header, 2), tracks = bigEndian(header, 2), ppq = bigEndian(header, 2);
    const bool timing = ppq != 0 && ppq < 32768;
    if (format > 1 || tracks == 0 || (format == 0 && tracks != 1) || !timing)
        reject("expected format 0/1 MIDI with PPQ timing");
    division = int(ppq);
    std::vector<Message> merged;
    for (uint32_t t = 0; t < tracks; ++t) {
        if (!chunkIs(file, "MTrk"))
            reject("missing MIDI track chunk");
        decodeTrack(chunk(file, bigEndian(file, 4)), merged);
    }
    std::stable_sort(merged.begin(), merged.end(), [](const Message& x, const Message& y) { return x.tick < y.tick; });
    return merged;
}

// QA inputs (src/io/midi.cpp:177-216): each completed note's onset frame,
// matched first-in first-out per channel and key, and the last note end.
void measureNotes(const std::vector<Message>& messages, int rate, int division, Midi& midi)
{
    std::vector<std::deque<int>> open(16 * 128);
    int quarter = 500000;
    uint64_t tick = 0;
    double seconds = 0;
    for (const auto& m : messages) {
        seconds += double(m.tick - tick) * (quarter * 1e-6 / division);
        tick = m.tick;
        if (m.kind == Message::tempo) {
            quarter = m.microseconds;
            continue;
        }
        if (!m.note())
            continue;
        if (seconds * rate >= INT_MAX / 2)
            reject("MIDI exceeds maximum render duration");
        const int frame = int(std::nearbyint(seconds * rate));
        auto& pending = open[size_t(m.channelIndex() * 128 + m.first)];
        if (m.sounding()) {
            pending.push_back(frame);
        } else if (!pending.empty()) {
            midi.onsets.push_back(pending.front());
            pending.pop_front();
            midi.bodyFrames = std::max(midi.bodyFrames, frame);
        }
    }
    std::sort(midi.onsets.begin(), midi.onsets.end());
    midi.onsets.erase(std::unique(midi.onsets.begin(), midi.onsets.end()), midi.onsets.end());
}

// Order at one tick once the natural ending is prepared: controllers, tempo
// and bends, then note ends, then note starts.
int tickOrder(const Message& m) { return m.sounding() ? 2 : m.noteEnd() ? 1 : 0; }

// Natural policy (src/io/midi.cpp:115-175). The render ends at the last
// completed note, or later where a sustain (64), sostenuto (66) or hold (69)
// pedal held at that point is released; everything after that end, every
// program change, the reference-level controllers 0/7/10/32/91/93 and panic
// commands (120/123) from the last completed note on are dropped. Volume 100
// and centre pan lead the job; the pedals are lifted at the end.
void naturalEnding(std::vector<Message>& messages)
{
    // Completed notes: a note end that closes a sounding note of that key.
    std::vector<int> sounding(16 * 128, 0);
    int channel = -1;
    uint64_t lastCompleted = 0;
    for (const auto& m : messages) {
        if (!m.note())
            continue;
        int& count = sounding[size_t(m.channelIndex() * 128 + m.first)];
        if (m.sounding()) {
            ++count;
            continue;
        }
        if (count == 0)
            continue;
        --count;
        if (channel >= 0 && channel != m.channelIndex())
            reject("multiple MIDI channels in one render job");
        channel = m.channelIndex();
        lastCompleted = m.tick;
    }
    if (channel < 0)
        reject("natural rendering requires a complete MIDI note");

    // Pedal releases after the last note move the end (one walk: an end moved
    // by one pedal changes which later messages count as "before the end").
    const auto pedalSlot = [](int cc) { return cc == 64 ? 0 : cc == 66 ? 1 : cc == 69 ? 2 : -1; };
    bool held[3] = {false, false, false};
    uint64_t end = lastCompleted;
    for (const auto& m : messages) {
        const int slot = m.type() == 0xb ? pedalSlot(m.first) : -1;
        if (slot < 0)
            continue;
        if (m.tick <= end)
            held[slot] = m.second >= 64;
        else if (held[slot] && m.second < 64) {
            end = m.tick;
            held[slot] = false;
        }
    }

    const auto dropped = [&](const Message& m) {
        if (m.kind == Message::trackEnd || m.tick > end || m.type() == 0xc)
            return true;
        if (m.type() != 0xb)
            return false;
        switch (m.first) {
        case 0: case 7: case 10: case 32: case 91: case 93:
            return true;
        case 120: case 123:
            return m.tick >= lastCompleted;
        default:
            return false;
        }
    };
    std::vector<Message> kept;
    kept.reserve(messages.size() + 6);
    const int control = 0xb0 | channel;
    kept.push_back(channelMessage(0, control, 7, 100));
    kept.push_back(channelMessage(0, control, 10, 64));
    for (const auto& m : messages)
        if (!dropped(m))
            kept.push_back(m);
    std::stable_sort(kept.begin(), kept.end(), [](const Message& x, const Message& y) {
        return x.tick != y.tick ? x.tick < y.tick : tickOrder(x) < tickOrder(y);
    });
    for (int cc : {64, 66, 69})
        kept.push_back(channelMessage(end, control, cc, 0));
    kept.push_back({Message::trackEnd, end, 0, 0, 0, 0});
    messages.swap(kept);
}

// Custom policy (src/io/midi.cpp:247-260): the job ends at its last note,
// controller or bend message, with the sustain pedal lifted there.
void customEnding(std::vector<Message>& messages)
{
    auto last = messages.end();
    for (auto it = messages.begin(); it != messages.end(); ++it) {
        const int type = it->type();
        if (type == 8 || type == 9 || type == 0xb || type == 0xe)
            last = it;
    }
    if (last == messages.end())
        return;
    const uint64_t tick = last->tick;
    const int channel = last->channelIndex();
    messages.erase(last + 1, messages.end());
    messages.push_back(channelMessage(tick, 0xb0 | channel, 64, 0));
    messages.push_back({Message::trackEnd, tick, 0, 0, 0, 0});
}

// Ticks to frames through the tempo changes, one message at a time
// (src/io/midi.cpp:262-274): natural rounds to nearest, the other policies
// floor with a 1e-7 guard.
class FrameClock {
public:
    FrameClock(int rate, int division, bool nearest) : scale_(rate / (1000000.0 * division)), nearest_(nearest) {}

    int at(uint64_t tick)
    {
        frames_ += double(tick - tick_) * quarter_ * scale_;
        tick_ = tick;
        if (frames_ >= INT_MAX / 2)
            reject("MIDI exceeds maximum render duration");
        return int(nearest_ ? std::nearbyint(frames_) : std::floor(frames_ + 1e-7));
    }
    void setTempo(int microseconds) { quarter_ = microseconds; }

private:
    double scale_, frames_ = 0;
    uint64_t tick_ = 0;
    int quarter_ = 500000;
    bool nearest_;
};

Midi parse(const std::string& data, int rate, TailPolicy tail, bool measure)
{
    int division = 0;
    std::vector<Message> messages = decodeTracks(data, division);
    Midi midi;
    if (measure)
        measureNotes(messages, rate, division, midi);
    if (tail == TailPolicy::natural)
        naturalEnding(messages);
    else if (tail == TailPolicy::custom)
        customEnding(messages);

    FrameClock clock(rate, division, tail == TailPolicy::natural);
    int channel = -1;
    for (const auto& m : messages) {
        const int frame = clock.at(m.tick);
        if (m.kind == Message::tempo) {
            clock.setTempo(m.microseconds);
            continue;
        }
        if (m.kind == Message::trackEnd) {
            midi.end = frame;  // the last end-of-track wins
            continue;
        }
        if (channel < 0)
            channel = m.channelIndex();
        else if (channel != m.channelIndex())
            reject("multiple MIDI channels in one render job");
        switch (m.type()) {
        case 8:
        case 9:
            midi.events.push_back({frame, m.sounding() ? EventKind::on : EventKind::off, m.first, m.second});
            break;
        case 0xb:
            midi.events.push_back({frame, EventKind::control, m.first, m.second});
            break;
        case 0xc:
            break;  // the job's SFZ is the preset
        case 0xe:
            midi.events.push_back({frame, EventKind::bend, (m.first | m.second << 7) - 8192, 0});
            break;
        default:
            reject("unsupported MIDI channel message type " + std::to_string(m.type()));
        }
    }
    return midi;
}

} // namespace

Midi parseMidi(const std::string& bytes, int rate, TailPolicy tail, bool measure)
{
    return parse(bytes, rate, tail, measure);
}

Midi loadMidi(const std::string& path, int rate, TailPolicy tail, bool measure)
{
    std::ifstream in(path, std::ios::binary);
    if (!in)
        throw std::runtime_error("cannot open MIDI " + path);
    const std::string data{std::istreambuf_iterator<char>(in), {}};
    try {
        return parse(data, rate, tail, measure);
    } catch (const std::exception& e) {
        throw std::runtime_error(path + ": " + e.what());
    }
}

} // namespace clay
