// SFZ loader. Grammar (headers, #include, #define, opcode=value with spaces
// in values), inheritance global > master > group > region, and the opcode
// meanings follow src/sfz/instrument.cpp of the reference.
#include "instrument.h"

#include <algorithm>
#include <cctype>
#include <cmath>
#include <filesystem>
#include <fstream>
#include <iterator>
#include <stdexcept>
#include <string_view>

namespace clay {
namespace {

namespace fs = std::filesystem;
using Opcodes = std::map<std::string, std::string>;
constexpr std::string_view kSpace = " \t\r\n";

std::string trimmed(std::string_view v)
{
    const auto a = v.find_first_not_of(kSpace);
    if (a == v.npos)
        return {};
    const auto b = v.find_last_not_of(kSpace);
    return std::string(v.substr(a, b - a + 1));
}

bool identChar(char c) { return std::isalnum(static_cast<unsigned char>(c)) || c == '_'; }

std::string unquoted(std::string v)
{
    if (v.size() >= 2 && v.front() == '"' && v.back() == '"')
        return v.substr(1, v.size() - 2);
    return v;
}

fs::path portablePath(std::string v)
{
    std::replace(v.begin(), v.end(), '\\', '/');
    return fs::u8path(v);
}

float real(const std::string& v, float lo, float hi)
{
    size_t used = 0;
    const float x = std::stof(v, &used);
    if (used != v.size() || !std::isfinite(x) || x < lo || x > hi)
        throw std::runtime_error("invalid value " + v);
    return x;
}

int64_t whole64(const std::string& v, int64_t lo, int64_t hi)
{
    size_t used = 0;
    const long long x = std::stoll(v, &used);
    if (used != v.size() || x < lo || x > hi)
        throw std::runtime_error("invalid value " + v);
    return x;
}

int whole(const std::string& v, int lo, int hi) { return int(whole64(v, lo, hi)); }

// MIDI key as a number or a note name such as c4 / F#3 / Bb-1 (C4 = 60).
int keyValue(const std::string& v)
{
    const auto letter = v.empty() ? 0 : std::toupper(static_cast<unsigned char>(v[0]));
    if (v.size() < 2 || letter < 'A' || letter > 'G')
        return whole(v, 0, 127);
    static const int semitone[7] = {9, 11, 0, 2, 4, 5, 7};
    int note = semitone[letter - 'A'];
    size_t at = 1;
    if (v[at] == '#' || v[at] == 'b')
        note += v[at++] == '#' ? 1 : -1;
    note += 12 * (whole(v.substr(at), -2, 10) + 1);
    if (note < 0 || note > 127)
        throw std::runtime_error("key outside MIDI range " + v);
    return note;
}

bool onOff(const std::string& v)
{
    if (v == "on" || v == "1")
        return true;
    if (v == "off" || v == "0")
        return false;
    throw std::runtime_error("invalid switch " + v);
}

bool curveIsPower(const std::string& v)
{
    if (v == "power")
        return true;
    if (v == "gain")
        return false;
    throw std::runtime_error("unsupported crossfade curve " + v);
}

bool isAmpEnvelope(ModTarget t) { return t >= ModTarget::envDelay && t <= ModTarget::envRelease; }

int controllerNumber(std::string_view digits)
{
    const int cc = whole(std::string(digits), 0, 511);
    if (cc >= 128 && cc <= 130)
        throw std::runtime_error("unsupported extended controller CC" + std::to_string(cc));
    return cc;
}

Modulation& bindingFor(std::vector<Modulation>& list, ModTarget target, int cc)
{
    auto it = std::find_if(list.begin(), list.end(), [&](const Modulation& m) { return m.target == target && m.cc == cc; });
    if (it == list.end()) {
        list.push_back({target, cc});
        return list.back();
    }
    return *it;
}

enum class Field { depth, curve, smooth, none };

// Split "_onccN" / "_ccN" / "ccN" / "_curveccN" / "_smoothccN" suffixes.
Field controllerField(std::string_view& suffix)
{
    const auto strip = [&](std::string_view prefix) {
        if (suffix.substr(0, prefix.size()) != prefix)
            return false;
        suffix.remove_prefix(prefix.size());
        return true;
    };
    if (strip("_oncc") || strip("_cc") || strip("cc"))
        return Field::depth;
    if (strip("_curvecc"))
        return Field::curve;
    if (strip("_smoothcc"))
        return Field::smooth;
    return Field::none;
}

// Controller modulation opcodes (src/sfz/instrument.cpp:114-194). With a
// nonzero envelope shift the ampeg_* names address a secondary envelope.
bool modulationOpcode(Region& r, const std::string& op, const std::string& value, int envelopeShift = 0,
                      std::array<float, 7>* velocityEnvelope = nullptr)
{
    static const std::pair<std::string_view, ModTarget> names[] = {
        {"amplitude", ModTarget::amplitude}, {"volume", ModTarget::volume}, {"gain", ModTarget::volume},
        {"pan", ModTarget::pan}, {"width", ModTarget::width}, {"position", ModTarget::position},
        {"pitch", ModTarget::pitch}, {"tune", ModTarget::pitch},
        {"ampeg_delay", ModTarget::envDelay}, {"ampeg_attack", ModTarget::envAttack},
        {"ampeg_hold", ModTarget::envHold}, {"ampeg_decay", ModTarget::envDecay},
        {"ampeg_sustain", ModTarget::envSustain}, {"ampeg_start", ModTarget::envStart},
        {"ampeg_release", ModTarget::envRelease}, {"offset", ModTarget::offset}, {"delay", ModTarget::delay},
        {"cutoff", ModTarget::cutoff}, {"cutoff1", ModTarget::cutoff}, {"cutoff2", ModTarget::cutoff2},
        {"resonance", ModTarget::resonance}, {"resonance1", ModTarget::resonance}, {"resonance2", ModTarget::resonance2},
        {"fil_veltrack", ModTarget::filterVelocity}, {"fil1_veltrack", ModTarget::filterVelocity},
        {"fil2_veltrack", ModTarget::filterVelocity2},
        {"eq1_freq", ModTarget::eq1Frequency}, {"eq1_bw", ModTarget::eq1Bandwidth}, {"eq1_gain", ModTarget::eq1Gain},
        {"eq2_freq", ModTarget::eq2Frequency}, {"eq2_bw", ModTarget::eq2Bandwidth}, {"eq2_gain", ModTarget::eq2Gain},
        {"eq3_freq", ModTarget::eq3Frequency}, {"eq3_bw", ModTarget::eq3Bandwidth}, {"eq3_gain", ModTarget::eq3Gain},
    };
    const std::string_view opcode(op);
    for (const auto& [name, base] : names) {
        if (envelopeShift && !isAmpEnvelope(base))
            continue;
        const auto target = static_cast<ModTarget>(int(base) + envelopeShift);
        if (isAmpEnvelope(base) && base != ModTarget::envStart) {
            const auto stage = std::string(name.substr(6));
            if (opcode == "ampeg_velto" + stage || opcode == "ampeg_vel2" + stage) {
                const float depth = real(value, -100, 100);
                auto& table = velocityEnvelope ? *velocityEnvelope : r.velocityEnvelope;
                table[size_t(int(base) - int(ModTarget::envDelay))] = base == ModTarget::envSustain ? depth / 100 : depth;
                return true;
            }
        }
        if (opcode.substr(0, name.size()) != name)
            continue;
        auto suffix = opcode.substr(name.size());
        const Field field = controllerField(suffix);
        if (field == Field::none)
            continue;
        if (target == ModTarget::offset && field == Field::curve) {
            (void)whole(std::string(suffix), 0, 511);
            (void)whole(value, 0, 255);
            return true;  // sfizz 1.2.3 ignores offset_curveccN
        }
        if ((field == Field::smooth && (envelopeShift || (target >= ModTarget::envDelay && target <= ModTarget::delay))) ||
            (field == Field::curve && target == ModTarget::delay))
            return false;
        auto& m = bindingFor(r.modulation, target, controllerNumber(suffix));
        if (field == Field::curve) {
            m.curve = whole(value, 0, 255);
        } else if (field == Field::smooth) {
            m.smooth = whole(value, 0, 10000);
        } else {
            const bool percent = target == ModTarget::amplitude || target == ModTarget::pan ||
                                 target == ModTarget::width || target == ModTarget::position ||
                                 base == ModTarget::envSustain || base == ModTarget::envStart;
            m.depth = real(value, -1e9f, 1e9f) / (percent ? 100 : 1);
        }
        if (target >= ModTarget::cutoff && target <= ModTarget::filterVelocity2) {
            const size_t slot = target == ModTarget::cutoff2 || target == ModTarget::resonance2 ||
                                target == ModTarget::filterVelocity2;
            if (r.filters[slot].type == FilterType::none)
                r.filters[slot].type = FilterType::low2;
        }
        if (target >= ModTarget::eq1Frequency)
            r.eq[size_t((int(target) - int(ModTarget::eq1Frequency)) / 3)].enabled = true;
        return true;
    }
    return false;
}

// fileg_* / pitcheg_* (src/sfz/instrument.cpp:196-236).
bool secondaryEnvelopeOpcode(Region& r, const std::string& op, const std::string& value)
{
    std::string_view name(op);
    int slot;
    if (name.substr(0, 6) == "fileg_") {
        slot = 0;
        name.remove_prefix(6);
    } else if (name.substr(0, 8) == "pitcheg_") {
        slot = 1;
        name.remove_prefix(8);
    } else {
        return false;
    }
    auto& env = r.secondary[size_t(slot)];
    env.enabled = true;
    if (name == "depth") {
        env.depth = real(value, -12000, 12000);
    } else if (name == "vel2depth" || name == "veltodepth") {
        env.velocityDepth = real(value, -12000, 12000);
    } else if (name == "dynamic") {
        env.dynamic = onOff(value);
    } else if ((slot == 0 && name.substr(0, 7) == "depthcc") || (slot == 1 && name == "decay_shape")) {
        return true;  // spellings unknown to sfizz 1.2.3
    } else if (name.substr(0, 10) == "depth_oncc" || name.substr(0, 13) == "depth_curvecc") {
        const bool curve = name.substr(0, 13) == "depth_curvecc";
        const int cc = controllerNumber(name.substr(curve ? 13 : 10));
        auto& m = bindingFor(r.modulation, slot == 0 ? ModTarget::filterEnvDepth : ModTarget::pitchEnvDepth, cc);
        if (curve)
            m.curve = whole(value, 0, 255);
        else
            m.depth = real(value, -12000, 12000);
    } else {
        static const std::string_view stages[] = {"delay", "attack", "hold", "decay", "sustain", "start", "release"};
        const auto found = std::find(std::begin(stages), std::end(stages), name);
        if (found != std::end(stages)) {
            const bool percent = name == "sustain" || name == "start";
            env.values[size_t(found - std::begin(stages))] = real(value, 0, 100) / (percent ? 100 : 1);
        } else {
            const auto first = slot == 0 ? ModTarget::filterEnvDelay : ModTarget::pitchEnvDelay;
            return modulationOpcode(r, "ampeg_" + std::string(name), value, int(first) - int(ModTarget::envDelay),
                                    &env.velocity);
        }
    }
    return true;
}

// lfoN_* and the legacy pitchlfo_/amplfo_/fillfo_ forms (src/sfz/instrument.cpp:238-324).
bool lfoOpcode(Region& r, const std::string& op, const std::string& value)
{
    std::string_view rest(op);
    int index;
    ModTarget legacy = ModTarget::pitch;
    if (rest.substr(0, 3) == "lfo") {
        rest.remove_prefix(3);
        const auto us = rest.find('_');
        if (us == rest.npos)
            return false;
        index = whole(std::string(rest.substr(0, us)), 1, 64);
        rest.remove_prefix(us + 1);
    } else if (rest.substr(0, 9) == "pitchlfo_") {
        index = -1;
        rest.remove_prefix(9);
    } else if (rest.substr(0, 7) == "amplfo_") {
        index = -2;
        legacy = ModTarget::volume;
        rest.remove_prefix(7);
    } else if (rest.substr(0, 7) == "fillfo_") {
        index = -3;
        legacy = ModTarget::cutoff;
        rest.remove_prefix(7);
    } else {
        return false;
    }
    if (rest.substr(0, 8) == "freq_lfo" && rest.find("_oncc") != rest.npos)
        return true;  // inter-LFO CC form unknown to sfizz 1.2.3
    auto& lfo = r.lfos[index];
    if (rest == "wave" || rest == "wave1") {
        const int wave = whole(value, -1, 12);
        if (wave >= 8 && wave < 12)
            throw std::runtime_error("unsupported LFO waveform " + value);
        lfo.wave = std::max(0, wave);
        return true;
    }
    static const std::pair<std::string_view, ModTarget> params[] = {
        {"freq", ModTarget::lfoFrequency}, {"phase", ModTarget::lfoPhase},
        {"delay", ModTarget::lfoDelay}, {"fade", ModTarget::lfoFade},
        {"pitch", ModTarget::pitch}, {"volume", ModTarget::volume}, {"pan", ModTarget::pan},
        {"width", ModTarget::width}, {"position", ModTarget::position},
        {"cutoff2", ModTarget::cutoff2}, {"cutoff1", ModTarget::cutoff}, {"cutoff", ModTarget::cutoff},
        {"resonance2", ModTarget::resonance2}, {"resonance1", ModTarget::resonance}, {"resonance", ModTarget::resonance},
        {"eq1freq", ModTarget::eq1Frequency}, {"eq1bw", ModTarget::eq1Bandwidth}, {"eq1gain", ModTarget::eq1Gain},
        {"eq2freq", ModTarget::eq2Frequency}, {"eq2bw", ModTarget::eq2Bandwidth}, {"eq2gain", ModTarget::eq2Gain},
        {"eq3freq", ModTarget::eq3Frequency}, {"eq3bw", ModTarget::eq3Bandwidth}, {"eq3gain", ModTarget::eq3Gain},
    };
    const auto assign = [&](std::string_view field, ModTarget target) {
        const float unit = target == ModTarget::pan || target == ModTarget::width || target == ModTarget::position ? 100.f : 1.f;
        if (field.empty()) {
            const float x = real(value, -24000, 24000) / unit;
            if (target == ModTarget::lfoFrequency) lfo.frequency = x;
            else if (target == ModTarget::lfoPhase) lfo.phase = x;
            else if (target == ModTarget::lfoDelay) lfo.delay = x;
            else if (target == ModTarget::lfoFade) lfo.fade = x;
            else lfo.depths[target] = x;
            return true;
        }
        const Field kind = controllerField(field);
        if (kind == Field::none)
            return false;
        auto& m = bindingFor(lfo.modulation, target, controllerNumber(field));
        if (kind == Field::curve) m.curve = whole(value, 0, 255);
        else if (kind == Field::smooth) m.smooth = whole(value, 0, 10000);
        else m.depth = real(value, -24000, 24000) / unit;
        return true;
    };
    if (index < 0 && rest.substr(0, 5) == "depth")
        return assign(rest.substr(5), legacy);
    for (const auto& [name, target] : params)
        if (rest.substr(0, name.size()) == name && assign(rest.substr(name.size()), target))
            return true;
    return false;
}

bool equalizerOpcode(Region& r, const std::string& op, const std::string& v)
{
    if (op.size() < 5 || op.compare(0, 2, "eq") != 0 || op[2] < '1' || op[2] > '3' || op[3] != '_')
        return false;
    auto& eq = r.eq[size_t(op[2] - '1')];
    const auto field = std::string_view(op).substr(4);
    if (field == "freq") eq.frequency = real(v, 0, 20000);
    else if (field == "gain") eq.gain = real(v, -96, 96);
    else if (field == "bw") eq.bandwidth = real(v, .001f, 12);
    else return false;
    eq.enabled = true;
    return true;
}

// egN_* flexible envelopes, pitch destination only (src/sfz/instrument.cpp:340-364).
bool flexEnvelopeOpcode(Region& r, const std::string& op, const std::string& v)
{
    if (op.compare(0, 2, "eg") != 0 || op.size() < 5 || op[2] < '0' || op[2] > '9')
        return false;
    const auto us = op.find('_', 2);
    if (us == op.npos)
        return false;
    auto& env = r.pitchEnvelopes[whole(op.substr(2, us - 2), 1, 64)];
    auto field = std::string_view(op).substr(us + 1);
    if (field == "sustain") {
        env.sustain = whole(v, 0, 63);
    } else if (field == "pitch") {
        env.depth = real(v, -12000, 12000);
    } else if (field.substr(0, 10) == "pitch_oncc") {
        env.depthCC[controllerNumber(field.substr(10))] = real(v, -12000, 12000);
    } else if (field.substr(0, 5) == "level") {
        env.points[whole(std::string(field.substr(5)), 0, 63)].level = real(v, -1, 1);
    } else if (field.substr(0, 4) == "time") {
        field.remove_prefix(4);
        const auto cc = field.find("_oncc");
        auto& point = env.points[whole(std::string(field.substr(0, cc)), 0, 63)];
        if (cc == field.npos)
            point.seconds = real(v, 0, 100);
        else
            point.timeCC[controllerNumber(field.substr(cc + 5))] = real(v, -100, 100);
    } else {
        return false;
    }
    return true;
}

bool filterOpcode(Region& r, const std::string& op, const std::string& v)
{
    std::string name = op;
    size_t slot = 0;
    if (name.compare(0, 5, "fil2_") == 0) {
        name.erase(3, 1);
        slot = 1;
    } else if (name.compare(0, 5, "fil1_") == 0) {
        name.erase(3, 1);
    } else if (name == "cutoff2" || name == "resonance2") {
        name.pop_back();
        slot = 1;
    } else if (name == "cutoff1" || name == "resonance1") {
        name.pop_back();
    }
    auto& f = r.filters[slot];
    if (name == "cutoff") f.cutoff = real(v, 0, 20000);
    else if (name == "resonance") f.resonance = real(v, -60, 96);
    else if (name == "fil_keytrack") f.keyTrack = real(v, -12000, 12000);
    else if (name == "fil_keycenter") f.keyCenter = keyValue(v);
    else if (name == "fil_veltrack") f.velocityTrack = real(v, -12000, 12000);
    else if (name == "fil_random" || name == "cutoff_random") f.random = real(v, -12000, 12000);
    else if (name == "fil_type") {
        static const std::map<std::string, FilterType> types{
            {"lpf_1p", FilterType::low1}, {"hpf_1p", FilterType::high1},
            {"lpf_2p", FilterType::low2}, {"hpf_2p", FilterType::high2},
            {"bpf_2p", FilterType::band2}, {"brf_2p", FilterType::notch2},
            {"lpf_4p", FilterType::low4}, {"hpf_4p", FilterType::high4}, {"bpf_4p", FilterType::band4},
            {"lpf_6p", FilterType::low6}, {"hpf_6p", FilterType::high6}, {"bpf_6p", FilterType::band6},
        };
        const auto it = types.find(v);
        if (it == types.end())
            throw std::runtime_error("unsupported filter type " + v);
        f.type = it->second;
    } else {
        return false;
    }
    if (f.type == FilterType::none)
        f.type = FilterType::low2;
    return true;
}

bool controllerFadeOpcode(Region& r, const std::string& op, const std::string& v)
{
    if (op == "xf_cccurve") {
        r.powerControllerFade = curveIsPower(v);
        return true;
    }
    static const std::pair<std::string_view, int Crossfade::*> fields[] = {
        {"xfin_locc", &Crossfade::inLow}, {"xfin_hicc", &Crossfade::inHigh},
        {"xfout_locc", &Crossfade::outLow}, {"xfout_hicc", &Crossfade::outHigh},
    };
    for (const auto& [prefix, member] : fields) {
        if (std::string_view(op).substr(0, prefix.size()) != prefix)
            continue;
        r.controllerFades[whole(op.substr(prefix.size()), 0, 127)].*member = whole(v, 0, 127);
        return true;
    }
    return false;
}

bool ignoredVariable(const std::string& op)
{
    if (op.compare(0, 3, "var") != 0)
        return false;
    const auto first = op.find_first_not_of("0123456789", 3);
    if (first == 3 || first == op.npos || op[first] != '_')
        return false;
    const std::string_view t(op.data() + first + 1, op.size() - first - 1);
    return t == "cutoff" || t == "mod" || t.substr(0, 4) == "oncc" || t.substr(0, 7) == "curvecc" || t.substr(0, 2) == "eq";
}

// Plain region opcodes (src/sfz/instrument.cpp:759-891).
void applyOpcode(Region& r, std::string& sample, std::map<int, Condition>& conditions, std::map<int, Condition>& triggers,
                 const std::string& op, const std::string& v)
{
    if (op == "sample") { sample = v; return; }
    if (op == "key" || op == "group_label" || op == "master_label" || op == "global_label" || op == "region_label" ||
        op == "sw_label")
        return;
    if (op == "lokey") { r.lowKey = v == "-1" ? 0 : keyValue(v); return; }
    if (op == "hikey") {
        r.triggerOnNote = v != "-1";
        r.highKey = r.triggerOnNote ? keyValue(v) : 127;
        return;
    }
    if (op == "pitch_keycenter") { r.rootKey = keyValue(v); return; }
    if (op == "lovel") { r.lowVelocity = whole(v, 0, 127); return; }
    if (op == "hivel") { r.highVelocity = whole(v, 0, 127); return; }
    if (op == "lochan" || op == "hichan") { (void)whole(v, 1, 16); return; }
    if (op == "volume") { r.volume += real(v, -144, 60); return; }
    if (op == "amplitude") { r.amplitude *= real(v, 0, 10000) / 100; return; }
    if (op == "amp_veltrack") { r.velocityTrack = real(v, -100, 100) / 100; return; }
    if (op == "amp_random") { r.ampRandom = real(v, -24, 24); return; }
    if (op == "pitch_random") { r.pitchRandom = real(v, -12000, 12000); return; }
    if (op.compare(0, 13, "amp_velcurve_") == 0) {
        const int velocity = whole(op.substr(13), 0, 127);
        const float value = real(v, 0, 1);
        auto it = std::find_if(r.velocityCurve.begin(), r.velocityCurve.end(), [&](const auto& p) { return p.first == velocity; });
        if (it == r.velocityCurve.end()) r.velocityCurve.emplace_back(velocity, value);
        else it->second = value;
        return;
    }
    if (op == "pan") { r.pan = real(v, -100, 100) / 100; return; }
    if (op == "width") { r.width = real(v, -100, 100) / 100; return; }
    if (op == "position") { r.position = real(v, -100, 100) / 100; return; }
    if (op == "xfin_lokey") { r.keyFade.inLow = keyValue(v); return; }
    if (op == "xfin_hikey") { r.keyFade.inHigh = keyValue(v); return; }
    if (op == "xfout_lokey") { r.keyFade.outLow = keyValue(v); return; }
    if (op == "xfout_hikey") { r.keyFade.outHigh = keyValue(v); return; }
    if (op == "xfin_lovel") { r.velocityFade.inLow = whole(v, 0, 127); return; }
    if (op == "xfin_hivel") { r.velocityFade.inHigh = whole(v, 0, 127); return; }
    if (op == "xfout_lovel") { r.velocityFade.outLow = whole(v, 0, 127); return; }
    if (op == "xfout_hivel") { r.velocityFade.outHigh = whole(v, 0, 127); return; }
    if (op == "xf_keycurve") { r.keyFade.power = curveIsPower(v); return; }
    if (op == "xf_velcurve") { r.velocityFade.power = curveIsPower(v); return; }
    if (op == "tune" || op == "pitch") { r.tune = real(v, -1200, 1200); return; }
    if (op == "transpose") { r.transpose = whole(v, -127, 127); return; }
    if (op == "pitch_keytrack") { r.keyTrack = real(v, -1200, 1200); return; }
    if (op == "bend_up" || op == "bendup") { r.bendUp = real(v, -12000, 12000); return; }
    if (op == "bend_down" || op == "benddown") { r.bendDown = real(v, -12000, 12000); return; }
    if (op == "bend_smooth") { r.bendSmooth = whole(v, 0, 10000); return; }
    if (op == "delay") { r.delay = real(v, 0, 100); return; }
    if (op == "group") { r.group = whole64(v, -INT32_MAX, UINT32_MAX); return; }
    if (op == "note_polyphony") { r.notePolyphony = whole(v, 1, 256); return; }
    if (op == "note_selfmask") { r.selfMask = v == "mask" || onOff(v); return; }
    if (op == "oscillator_phase") { r.oscillatorPhase = real(v, -1000, 1000); return; }
    if (op == "off_by" || op == "offby") {
        r.offBy = whole64(v, -INT32_MAX, UINT32_MAX);
        r.chokedByGroup = true;
        return;
    }
    if (op == "off_mode" || op == "offmode") {
        if (v == "fast") r.offMode = OffMode::fast;
        else if (v == "normal") r.offMode = OffMode::normal;
        else if (v == "time") r.offMode = OffMode::time;
        else throw std::runtime_error("unsupported off mode " + v);
        return;
    }
    if (op == "off_time") {
        r.offTime = real(v, 0, 100);
        r.offMode = OffMode::time;
        return;
    }
    if (op == "offset") { r.offset = whole(v, 0, INT32_MAX - 1); return; }
    if (op == "offset_random") { r.offsetRandom = whole(v, 0, INT32_MAX - 1); return; }
    if (op == "end") { r.end = whole(v, 0, INT32_MAX - 1); return; }
    if (op == "loop_start" || op == "loopstart") { r.loopStart = whole(v, 0, INT32_MAX - 1); return; }
    if (op == "loop_end" || op == "loopend") { r.loopEnd = whole(v, 0, INT32_MAX - 1); return; }
    if (op == "loop_crossfade") { r.loopCrossfade = std::max(.001f, real(v, 0, 100)); return; }
    if (op == "loop_type" || op == "looptype") {
        if (v != "forward" && v != "alternate" && v != "backward")
            throw std::runtime_error("unknown loop direction hint " + v);
        return;  // no effect in sfizz 1.2.3
    }
    if (op == "ampeg_delay") { r.envelopeDelay = real(v, 0, 100); return; }
    if (op == "ampeg_attack") { r.attack = real(v, 0, 100); return; }
    if (op == "ampeg_hold") { r.hold = real(v, 0, 100); return; }
    if (op == "ampeg_decay") { r.decay = real(v, 0, 100); return; }
    if (op == "ampeg_sustain") { r.sustain = real(v, 0, 100) / 100; return; }
    if (op == "ampeg_start") { r.start = real(v, 0, 100) / 100; return; }
    if (op == "ampeg_release") { r.release = real(v, 0, 100); return; }
    if (op == "ampeg_dynamic") { r.envelopeDynamic = onOff(v); return; }
    if (op == "rt_decay") { r.releaseDecay = real(v, 0, 200); return; }
    if (op == "rt_dead") { r.releaseDead = onOff(v); return; }
    if (op == "sustain_sw") { r.sustainEnabled = onOff(v); return; }
    if (op == "sustain_cc") { r.sustainCC = whole(v, 0, 127); return; }
    if (op == "sustain_lo") { r.sustainLow = whole(v, 0, 127); return; }
    if (op == "seq_length") { r.sequenceLength = whole(v, 1, 1000000); return; }
    if (op == "seq_position") {
        r.sequencePosition = whole(v, 1, 1000000);
        r.explicitSequencePosition = true;
        return;
    }
    if (op == "sw_default") { r.switchDefault = keyValue(v); return; }
    if (op == "sw_down") { r.switchDown = keyValue(v); return; }
    if (op == "sw_up") { r.switchUp = keyValue(v); return; }
    if (op == "sw_previous") { r.switchPrevious = keyValue(v); return; }
    if (op == "sw_lokey" || op == "sw_hikey") { (void)keyValue(v); return; }
    if (op == "sw_last") {
        if (!r.switchRange)
            r.switchLow = r.switchHigh = keyValue(v);
        return;
    }
    if (op == "sw_lolast" || op == "sw_hilast") {
        const int note = keyValue(v);
        if (!r.switchRange)
            r.switchLow = r.switchHigh = note;
        (op == "sw_lolast" ? r.switchLow : r.switchHigh) = note;
        r.switchRange = true;
        return;
    }
    if (op == "lorand") { r.randomLow = real(v, 0, 1); return; }
    if (op == "hirand") { r.randomHigh = real(v, 0, 1); return; }
    if (op == "trigger") {
        static const std::map<std::string, Trigger> values{{"attack", Trigger::attack}, {"release", Trigger::release},
            {"release_key", Trigger::releaseKey}, {"first", Trigger::first}, {"legato", Trigger::legato}};
        r.trigger = values.at(v);
        return;
    }
    if (op == "loop_mode") {
        static const std::map<std::string, Loop> values{{"no_loop", Loop::none}, {"one_shot", Loop::oneShot},
            {"loop_continuous", Loop::continuous}, {"loop_sustain", Loop::sustain}};
        r.loop = values.at(v);
        return;
    }
    if (modulationOpcode(r, op, v) || filterOpcode(r, op, v) || equalizerOpcode(r, op, v) || lfoOpcode(r, op, v) ||
        secondaryEnvelopeOpcode(r, op, v) || flexEnvelopeOpcode(r, op, v) || controllerFadeOpcode(r, op, v) ||
        ignoredVariable(op) || op == "ampeg_attack_shape" || op == "ampeg_release_shape")
        return;
    const bool onTrigger = op.compare(0, 3, "on_") == 0;
    const size_t at = onTrigger ? 3 : 0;
    if (op.compare(at, 4, "locc") != 0 && op.compare(at, 4, "hicc") != 0)
        throw std::runtime_error("unsupported SFZ opcode");
    const auto cc = uint8_t(whole(op.substr(at + 4), 0, 127));
    auto& ranges = onTrigger ? triggers : conditions;
    auto& c = ranges.emplace(cc, Condition{cc, 0, 127}).first->second;
    (op[at] == 'l' ? c.low : c.high) = uint8_t(whole(v, 0, 127));
}

class Loader {
public:
    Loader(fs::path root, AudioBank& bank) : root_(std::move(root)), bank_(bank) {}

    Instrument load(const fs::path& path)
    {
        parseFile(path, 0);
        closeBlock();
        if (out_.regions.empty())
            throw std::runtime_error("no playable SFZ regions in " + path.u8string());
        return std::move(out_);
    }

private:
    void parseFile(const fs::path& file, int depth)
    {
        if (depth >= 32)
            throw std::runtime_error("too many nested SFZ includes");
        std::ifstream in(file, std::ios::binary);
        if (!in)
            throw std::runtime_error("cannot open SFZ " + file.u8string());
        const std::string text{std::istreambuf_iterator<char>(in), {}};
        const std::string_view src(text);
        size_t pos = src.substr(0, 3) == "\xef\xbb\xbf" ? 3 : 0;
        try {
            while (pos < src.size()) {
                pos = src.find_first_not_of(kSpace, pos);
                if (pos == src.npos)
                    break;
                if (src.substr(pos, 2) == "//") {
                    pos = src.find('\n', pos);
                } else if (src[pos] == '<') {
                    const auto close = src.find('>', pos + 1);
                    if (close == src.npos)
                        throw std::runtime_error("unclosed SFZ header");
                    openHeader(trimmed(src.substr(pos + 1, close - pos - 1)));
                    pos = close + 1;
                } else if (src[pos] == '#') {
                    const auto eol = src.find('\n', pos);
                    directive(file, src.substr(pos, eol - pos), depth);
                    pos = eol;
                } else {
                    pos = opcodeAt(src, pos);
                }
            }
        } catch (const std::exception& e) {
            const auto line = 1 + std::count(src.begin(), src.begin() + std::min(pos, src.size()), '\n');
            throw std::runtime_error(file.u8string() + ":" + std::to_string(line) + ": " + e.what());
        }
    }

    void directive(const fs::path& file, std::string_view text, int depth)
    {
        if (text.substr(0, 8) == "#include") {
            const auto name = trimmed(text.substr(8));
            if (name.size() < 2 || name.front() != '"' || name.back() != '"')
                throw std::runtime_error("expected quoted #include path");
            parseFile(file.parent_path() / portablePath(unquoted(expand(name))), depth + 1);
        } else if (text.substr(0, 7) == "#define") {
            const auto def = trimmed(text.substr(7));
            const auto gap = def.find_first_of(kSpace);
            if (def.empty() || def.front() != '$' || gap == def.npos)
                throw std::runtime_error("invalid #define");
            macros_[def.substr(0, gap)] = expand(trimmed(std::string_view(def).substr(gap)));
        } else {
            throw std::runtime_error("unsupported SFZ directive " + std::string(text));
        }
    }

    // Parse `name=value` at pos; a value runs to end of line, `<`, `//`, or the
    // next `name=` after whitespace (quotes protect). Returns the new position.
    size_t opcodeAt(std::string_view src, size_t pos)
    {
        size_t end = pos;
        while (end < src.size() && (identChar(src[end]) || src[end] == '$'))
            ++end;
        const auto eq = src.find_first_not_of(" \t", end);
        if (end == pos || eq == src.npos || src[eq] != '=')
            throw std::runtime_error("expected opcode=value");
        const auto name = expand(std::string(src.substr(pos, end - pos)));
        size_t at = eq + 1;
        bool quoted = false;
        while (at < src.size()) {
            const char c = src[at];
            if (c == '"')
                quoted = !quoted;
            if (!quoted && (c == '\r' || c == '\n' || c == '<' || src.substr(at, 2) == "//"))
                break;
            if (!quoted && (c == ' ' || c == '\t')) {
                const auto next = src.find_first_not_of(" \t", at);
                auto word = next;
                while (word < src.size() && (identChar(src[word]) || src[word] == '$'))
                    ++word;
                const auto sep = src.find_first_not_of(" \t", word);
                if (word != next && sep < src.size() && src[sep] == '=')
                    break;
            }
            ++at;
        }
        if (quoted)
            throw std::runtime_error("unclosed quoted value for " + name);
        opcode(name, unquoted(expand(trimmed(src.substr(eq + 1, at - eq - 1)))));
        return at;
    }

    std::string expand(std::string v) const
    {
        size_t at = 0;
        while ((at = v.find('$', at)) != v.npos) {
            size_t end = at + 1;
            while (end < v.size() && identChar(v[end]))
                ++end;
            const auto name = v.substr(at, end - at);
            const auto it = macros_.find(name);
            if (it == macros_.end())
                throw std::runtime_error("undefined SFZ macro " + name);
            v.replace(at, end - at, it->second);
            at += it->second.size();
        }
        return v;
    }

    void closeBlock()
    {
        if (header_ == "curve")
            finishCurve();
        else
            finishRegion();
    }

    void openHeader(const std::string& name)
    {
        closeBlock();
        header_ = name;
        if (name == "global") {
            global_.clear();
            master_.clear();
            group_.clear();
        } else if (name == "master") {
            master_.clear();
            group_.clear();
        } else if (name == "group") {
            group_.clear();
        } else if (name != "region" && name != "control" && name != "curve") {
            throw std::runtime_error("unsupported SFZ header <" + name + ">");
        }
    }

    void opcode(const std::string& name, const std::string& value)
    {
        if (value.empty())
            throw std::runtime_error("empty value for " + name);
        if (header_ == "global") global_[name] = value;
        else if (header_ == "master") master_[name] = value;
        else if (header_ == "group") group_[name] = value;
        else if (header_ == "region") region_[name] = value;
        else if (header_ == "curve") curve_[name] = value;
        else if (header_ == "control") controlOpcode(name, value);
        else throw std::runtime_error("opcode outside SFZ header " + name);
    }

    void controlOpcode(const std::string& name, const std::string& value)
    {
        if (name == "default_path") {
            defaultPath_ = portablePath(value);
        } else if (name.compare(0, 6, "set_cc") == 0 || name.compare(0, 8, "set_hdcc") == 0) {
            const bool hd = name[4] == 'h';
            const int cc = whole(name.substr(hd ? 8 : 6), 0, 511);
            const float raw = hd ? 127 * real(value, 0, 1) : real(value, 0, 127);
            if (cc < 128) out_.initialCC[size_t(cc)] = raw;
            else out_.extendedCCDefaults[cc] = raw;
        } else if (name.compare(0, 8, "label_cc") == 0) {
            (void)whole(name.substr(8), 0, 511);
        } else if (name.compare(0, 9, "label_key") == 0) {
            (void)whole(name.substr(9), 0, 127);
        } else if (name != "image" && name != "image_controls") {
            throw std::runtime_error("unsupported control opcode " + name);
        }
    }

    void finishCurve()
    {
        const auto index = curve_.find("curve_index");
        if (index == curve_.end() && explicitCurves_) {
            curve_.clear();
            return;
        }
        const int id = index == curve_.end() ? nextCurve_++ : whole(index->second, 0, 255);
        explicitCurves_ |= index != curve_.end();
        std::map<int, float> points{{0, 0.f}, {127, 1.f}};
        for (const auto& [op, value] : curve_) {
            if (op == "curve_index")
                continue;
            if (op.front() != 'v')
                throw std::runtime_error("unsupported <curve> opcode " + op);
            points[whole(op.substr(1), 0, 127)] = real(value, -1e16f, 1e16f);
        }
        auto& curve = out_.curves[id];
        for (auto a = points.begin(), b = std::next(a); b != points.end(); ++a, ++b)
            for (int x = a->first; x <= b->first; ++x) {
                const float t = float(x - a->first) / float(b->first - a->first);
                curve[size_t(x)] = std::clamp(a->second + t * (b->second - a->second), -1.f, 1.f);
            }
        curve_.clear();
    }

    void finishRegion()
    {
        if (header_ != "region")
            return;
        Region r;
        std::string sample;
        std::map<int, Condition> conditions, triggers;
        for (const Opcodes* scope : {&global_, &master_, &group_, &region_}) {
            if (const auto it = scope->find("key"); it != scope->end()) {
                r.triggerOnNote = it->second != "-1";
                r.lowKey = r.highKey = r.rootKey = r.triggerOnNote ? keyValue(it->second) : 60;
            }
            for (const auto& [op, value] : *scope) {
                try {
                    applyOpcode(r, sample, conditions, triggers, op, value);
                } catch (const std::exception& e) {
                    throw std::runtime_error(op + "=" + value + ": " + e.what());
                }
            }
        }
        const bool emptyBlock = region_.empty();
        region_.clear();
        if (sample.empty() && emptyBlock)
            sample = "*silence";
        if (sample == "*sine") r.generator = Generator::sine;
        else if (sample == "*silence") r.generator = Generator::silence;
        else if (sample == "*noise") r.generator = Generator::noise;
        else if (sample.empty() || sample.front() == '*')
            throw std::runtime_error("missing or unsupported region sample " + sample);
        if ((r.lowKey > r.highKey && triggers.empty()) || r.lowVelocity > r.highVelocity || r.randomLow > r.randomHigh ||
            r.switchLow > r.switchHigh)
            throw std::runtime_error("invalid region range for " + sample);
        if (!r.velocityCurve.empty()) {
            std::sort(r.velocityCurve.begin(), r.velocityCurve.end());
            if (r.velocityCurve.front().first != 0)
                r.velocityCurve.insert(r.velocityCurve.begin(), {0, 0.f});
            if (r.velocityCurve.back().first != 127)
                r.velocityCurve.emplace_back(127, 1.f);
        }
        if (r.generator == Generator::none) {
            r.sample = bank_.describe((root_ / defaultPath_ / portablePath(sample)).string());
            const Sample& audio = bank_.samples[size_t(r.sample)];
            r.end = r.end < 0 ? audio.frames : std::min(r.end + 1, audio.frames);
            if (r.offset >= r.end)
                throw std::runtime_error("offset past end of sample " + sample);
            if (r.loop == Loop::automatic)
                r.loop = audio.loopEnd > 0 || r.loopEnd >= 0 ? Loop::continuous : Loop::none;
            if (r.loop == Loop::continuous || r.loop == Loop::sustain) {
                if (r.loopStart < 0) r.loopStart = audio.loopFirst;
                if (r.loopEnd < 0) r.loopEnd = audio.loopEnd > 0 ? audio.loopEnd : r.end;
                else r.loopEnd = std::min(r.loopEnd + 1, r.end);
                if (r.loopStart >= r.loopEnd || r.loopEnd > r.end || r.offset >= r.loopEnd)
                    throw std::runtime_error("invalid forward loop in " + sample);
            }
        }
        for (const auto& [cc, c] : conditions) {
            if (c.low > c.high)
                throw std::runtime_error("invalid CC range for " + sample);
            r.conditions.push_back(c);
            out_.usedCC[size_t(cc)] = true;
        }
        for (auto& [cc, fade] : r.controllerFades) {
            fade.power = r.powerControllerFade;
            out_.usedCC[size_t(cc)] = true;
        }
        for (const auto& m : r.modulation)
            if (m.cc < 128)
                out_.usedCC[size_t(m.cc)] = true;
        for (const auto& [id, lfo] : r.lfos) {
            const auto enable = [&](ModTarget t) {
                const size_t second = t == ModTarget::cutoff2 || t == ModTarget::resonance2;
                if (t >= ModTarget::cutoff && t <= ModTarget::resonance2 && r.filters[second].type == FilterType::none)
                    r.filters[second].type = FilterType::low2;
                if (t >= ModTarget::eq1Frequency)
                    r.eq[size_t((int(t) - int(ModTarget::eq1Frequency)) / 3)].enabled = true;
            };
            for (const auto& depth : lfo.depths)
                enable(depth.first);
            for (const auto& m : lfo.modulation) {
                if (m.cc < 128)
                    out_.usedCC[size_t(m.cc)] = true;
                enable(m.target);
            }
        }
        for (const auto& [id, env] : r.pitchEnvelopes) {
            if (env.points.empty() || env.points.rbegin()->second.level != 0 ||
                (env.sustain >= 0 && env.sustain != env.points.rbegin()->first))
                throw std::runtime_error("flexible pitch envelope must end at a zero-level sustain point");
            for (const auto& cc : env.depthCC)
                if (cc.first < 128)
                    out_.usedCC[size_t(cc.first)] = true;
            for (const auto& point : env.points)
                for (const auto& cc : point.second.timeCC)
                    if (cc.first < 128)
                        out_.usedCC[size_t(cc.first)] = true;
        }
        if (r.sustainEnabled)
            out_.usedCC[size_t(r.sustainCC)] = true;
        const int index = int(out_.regions.size());
        if (r.triggerOnNote)
            for (int note = r.lowKey; note <= r.highKey; ++note)
                out_.byKey[size_t(note)].push_back(index);
        for (const auto& [cc, c] : triggers) {
            if (c.low > c.high)
                throw std::runtime_error("invalid CC trigger range for " + sample);
            r.ccTriggers.push_back(c);
            out_.byCC[size_t(cc)].push_back(index);
            out_.usedCC[size_t(cc)] = true;
        }
        if (r.switchLow >= 0)
            for (int note = r.switchLow; note <= r.switchHigh; ++note)
                out_.switchKeys[size_t(note)] = true;
        if (r.switchDefault >= 0)
            out_.initialSwitch = r.switchDefault;
        out_.hasOffBy |= r.chokedByGroup;
        out_.regions.push_back(std::move(r));
    }

    fs::path root_, defaultPath_;
    AudioBank& bank_;
    Instrument out_;
    std::string header_;
    Opcodes global_, master_, group_, region_, macros_, curve_;
    int nextCurve_ = 7;
    bool explicitCurves_ = false;
};

} // namespace

Instrument loadInstrument(const std::string& path, AudioBank& bank)
{
    const fs::path p(path);
    return Loader(p.parent_path(), bank).load(p);
}

float curveValue(const Instrument& instrument, int curve, float value)
{
    value = std::clamp(value, 0.f, 127.f);
    const int lo = int(value);
    const int hi = std::min(127, lo + 1);
    const float t = value - float(lo);
    if (const auto it = instrument.curves.find(curve); it != instrument.curves.end())
        return it->second[size_t(lo)] + t * (it->second[size_t(hi)] - it->second[size_t(lo)]);
    const auto at = [curve](int k) {
        const float x = float(k) / 127.f;
        switch (curve) {
        case 1: return 2 * x - 1;
        case 2: return 1 - x;
        case 3: return 1 - 2 * x;
        case 4: return x * x;
        case 5: return std::sqrt(x);
        case 6: return std::sqrt(1 - x);
        default: return x;
        }
    };
    const float a = at(lo);
    return a + t * (at(hi) - a);
}

} // namespace clay
