// SFZ loader: text -> scoped opcode sets -> regions.
//
// The file is read in two layers. SfzReader turns the text (with #include
// and #define) into header and opcode events, following the SFZ syntax rules
// sfizz applies (src/sfz/instrument.cpp:455-533): a value runs to the end of
// the line, a header, a comment or the next `name=`. RegionBuilder keeps the
// opcodes of the open <global>, <master>, <group> and <region> scopes and,
// when a region closes, applies them in scope order (and by name within a
// scope, as a later definition of the same opcode replaces an earlier one),
// then resolves the sample, its loop and the region's lookup tables
// (:634-757). Opcode units and ranges are sfizz's (:759-891).
#include "instrument.h"

#include <algorithm>
#include <cctype>
#include <cmath>
#include <fstream>
#include <functional>
#include <iterator>
#include <map>
#include <stdexcept>
#include <string_view>

namespace clay {
namespace {

using Scope = std::map<std::string, std::string>;

[[noreturn]] void bad(const std::string& why) { throw std::runtime_error(why); }

bool isNameChar(char c) { return std::isalnum(static_cast<unsigned char>(c)) || c == '_'; }

std::string trimmed(std::string_view s)
{
    const auto first = s.find_first_not_of(" \t\r\n");
    if (first == s.npos)
        return {};
    return std::string(s.substr(first, s.find_last_not_of(" \t\r\n") - first + 1));
}

std::string withoutQuotes(const std::string& s)
{
    return s.size() >= 2 && s.front() == '"' && s.back() == '"' ? s.substr(1, s.size() - 2) : s;
}

bool startsWith(std::string_view s, std::string_view prefix) { return s.substr(0, prefix.size()) == prefix; }

// Numbers must use the whole text and lie in range.
float real(const std::string& text, float low, float high)
{
    size_t used = 0;
    const float v = std::stof(text, &used);
    if (used != text.size() || !std::isfinite(v) || v < low || v > high)
        bad("invalid value " + text);
    return v;
}

int64_t whole64(const std::string& text, int64_t low, int64_t high)
{
    size_t used = 0;
    const long long v = std::stoll(text, &used);
    if (used != text.size() || v < low || v > high)
        bad("invalid value " + text);
    return v;
}

int whole(const std::string& text, int low, int high) { return int(whole64(text, low, high)); }

// A MIDI key as a number or a note name with octave (C4 = 60).
int midiKey(const std::string& text)
{
    const char letter = char(std::toupper(static_cast<unsigned char>(text.empty() ? '\0' : text[0])));
    if (text.size() < 2 || letter < 'A' || letter > 'G')
        return whole(text, 0, 127);
    static const int semitone[7] = {9, 11, 0, 2, 4, 5, 7};  // A B C D E F G
    int note = semitone[letter - 'A'];
    size_t digits = 1;
    if (text[1] == '#')
        ++note, ++digits;
    else if (text[1] == 'b')
        --note, ++digits;
    note += 12 * (whole(text.substr(digits), -2, 10) + 1);
    if (note < 0 || note > 127)
        bad("key outside MIDI range " + text);
    return note;
}

bool onOff(const std::string& text)
{
    if (text == "on" || text == "1")
        return true;
    if (text == "off" || text == "0")
        return false;
    bad("invalid switch " + text);
}

std::string slashed(std::string path)
{
    std::replace(path.begin(), path.end(), '\\', '/');
    return path;
}

std::string joinPath(const std::string& a, const std::string& b)
{
    if (a.empty() || (!b.empty() && b.front() == '/'))
        return b;
    return a.back() == '/' ? a + b : a + "/" + b;
}

std::string directoryOf(const std::string& path)
{
    const auto slash = path.find_last_of('/');
    return slash == std::string::npos ? std::string() : path.substr(0, slash);
}

// Controller sources a binding may name: CC 0..127, and extended sources
// except 128..130 (sfizz 1.2.3) and the random/alternating ones 135..137,
// which this render core does not model.
int source(std::string_view digits)
{
    const int cc = whole(std::string(digits), 0, 511);
    if ((cc >= 128 && cc <= 130) || (cc >= 135 && cc <= 137))
        bad("unsupported extended controller CC" + std::to_string(cc));
    return cc;
}

// Parameter names that take controller bindings, with their targets
// (src/sfz/instrument.cpp:113-127); searched in this order by prefix.
struct Parameter {
    std::string_view name;
    ModTarget target;
};
constexpr Parameter kParameters[] = {
    {"amplitude", ModTarget::amplitude}, {"volume", ModTarget::volume}, {"gain", ModTarget::volume},
    {"pan", ModTarget::pan}, {"width", ModTarget::width}, {"position", ModTarget::position},
    {"pitch", ModTarget::pitch}, {"tune", ModTarget::pitch},
    {"ampeg_delay", ModTarget::envDelay}, {"ampeg_attack", ModTarget::envAttack}, {"ampeg_hold", ModTarget::envHold},
    {"ampeg_decay", ModTarget::envDecay}, {"ampeg_sustain", ModTarget::envSustain},
    {"ampeg_start", ModTarget::envStart}, {"ampeg_release", ModTarget::envRelease},
    {"offset", ModTarget::offset}, {"delay", ModTarget::delay},
    {"cutoff", ModTarget::cutoff}, {"cutoff1", ModTarget::cutoff}, {"cutoff2", ModTarget::cutoff2},
    {"resonance", ModTarget::resonance}, {"resonance1", ModTarget::resonance},
    {"resonance2", ModTarget::resonance2},
    {"fil_veltrack", ModTarget::filterVelocity}, {"fil1_veltrack", ModTarget::filterVelocity},
    {"fil2_veltrack", ModTarget::filterVelocity2},
    {"eq1_freq", ModTarget::eq1Frequency}, {"eq1_bw", ModTarget::eq1Bandwidth}, {"eq1_gain", ModTarget::eq1Gain},
    {"eq2_freq", ModTarget::eq2Frequency}, {"eq2_bw", ModTarget::eq2Bandwidth}, {"eq2_gain", ModTarget::eq2Gain},
    {"eq3_freq", ModTarget::eq3Frequency}, {"eq3_bw", ModTarget::eq3Bandwidth}, {"eq3_gain", ModTarget::eq3Gain},
};

bool ampStage(ModTarget t) { return t >= ModTarget::envDelay && t <= ModTarget::envRelease; }

enum class BindingField { depth, curve, smooth, none };

// "_onccN", "_ccN", "ccN" (depth), "_curveccN", "_smoothccN" after a parameter name.
BindingField bindingField(std::string_view& rest)
{
    for (const auto& [tag, field] : {std::pair<std::string_view, BindingField>{"_oncc", BindingField::depth},
                                     {"_cc", BindingField::depth}, {"cc", BindingField::depth},
                                     {"_curvecc", BindingField::curve}, {"_smoothcc", BindingField::smooth}})
        if (startsWith(rest, tag)) {
            rest.remove_prefix(tag.size());
            return field;
        }
    return BindingField::none;
}

Binding& bindingFor(Region& r, ModTarget target, int cc)
{
    for (auto& b : r.bindings)
        if (b.target == target && b.source == cc)
            return b;
    r.bindings.push_back({target, cc});
    return r.bindings.back();
}

void enableFilterOf(Region& r, ModTarget t)
{
    if (t >= ModTarget::cutoff && t <= ModTarget::filterVelocity2) {
        const bool second = t == ModTarget::cutoff2 || t == ModTarget::resonance2 || t == ModTarget::filterVelocity2;
        if (r.filters[second].type == FilterType::none)
            r.filters[second].type = FilterType::low2;
    }
    if (t >= ModTarget::eq1Frequency)
        r.eq[size_t((int(t) - int(ModTarget::eq1Frequency)) / 3)].used = true;
}

// Controller bindings and amplitude-envelope velocity offsets
// (src/sfz/instrument.cpp:111-193). `shift` moves amplitude-envelope targets
// to a filter or pitch envelope's; `velocity` receives envelope velocity
// offsets.
bool bindingOpcode(Region& r, const std::string& op, const std::string& value, int shift, Stages& velocity)
{
    const std::string_view name(op);
    for (const auto& p : kParameters) {
        if (shift && !ampStage(p.target))
            continue;
        const auto target = ModTarget(int(p.target) + shift);
        if (ampStage(p.target) && p.target != ModTarget::envStart) {
            const std::string_view stage = p.name.substr(6);  // after "ampeg_"
            if (name == "ampeg_velto" + std::string(stage) || name == "ampeg_vel2" + std::string(stage)) {
                const float offset = real(value, -100, 100);
                velocity[size_t(int(p.target) - int(ModTarget::envDelay))] =
                    p.target == ModTarget::envSustain ? offset / 100 : offset;
                return true;
            }
        }
        if (!startsWith(name, p.name))
            continue;
        std::string_view rest = name.substr(p.name.size());
        const BindingField field = bindingField(rest);
        if (field == BindingField::none)
            continue;
        if (target == ModTarget::offset && field == BindingField::curve) {
            (void)whole(std::string(rest), 0, 511);  // accepted and ignored, as sfizz 1.2.3 does
            (void)whole(value, 0, 255);
            return true;
        }
        const bool noSmooth = shift || (target >= ModTarget::envDelay && target <= ModTarget::delay);
        if ((field == BindingField::smooth && noSmooth) || (field == BindingField::curve && target == ModTarget::delay))
            return false;
        Binding& b = bindingFor(r, target, source(rest));
        if (field == BindingField::curve)
            b.curve = whole(value, 0, 255);
        else if (field == BindingField::smooth)
            b.smoothMs = whole(value, 0, 10000);
        else {
            const bool percent = target == ModTarget::amplitude || target == ModTarget::pan ||
                                 target == ModTarget::width || target == ModTarget::position ||
                                 p.target == ModTarget::envSustain || p.target == ModTarget::envStart;
            b.depth = real(value, -1e9f, 1e9f) / (percent ? 100 : 1);
        }
        enableFilterOf(r, target);
        return true;
    }
    return false;
}

// fileg_* and pitcheg_* (src/sfz/instrument.cpp:196-234).
bool modEnvelopeOpcode(Region& r, const std::string& op, const std::string& value)
{
    int slot;
    std::string_view rest(op);
    if (startsWith(rest, "fileg_"))
        slot = 0, rest.remove_prefix(6);
    else if (startsWith(rest, "pitcheg_"))
        slot = 1, rest.remove_prefix(8);
    else
        return false;
    ModEnvelope& env = r.modEnvelopes[size_t(slot)];
    env.used = true;
    static constexpr std::string_view kStages[stageCount] = {"delay", "attack", "hold", "decay", "sustain", "start",
                                                             "release"};
    if (rest == "depth")
        env.depth = real(value, -12000, 12000);
    else if (rest == "vel2depth" || rest == "veltodepth")
        env.velocityDepth = real(value, -12000, 12000);
    else if ((slot == 0 && startsWith(rest, "depthcc")) || (slot == 1 && rest == "decay_shape"))
        return true;  // spellings sfizz 1.2.3 does not know: no effect
    else if (startsWith(rest, "depth_oncc") || startsWith(rest, "depth_curvecc")) {
        const bool curve = startsWith(rest, "depth_curvecc");
        Binding& b = bindingFor(r, slot == 0 ? ModTarget::filterEnvDepth : ModTarget::pitchEnvDepth,
                                source(rest.substr(curve ? 13 : 10)));
        if (curve)
            b.curve = whole(value, 0, 255);
        else
            b.depth = real(value, -12000, 12000);
    } else if (const auto* at = std::find(std::begin(kStages), std::end(kStages), rest); at != std::end(kStages)) {
        const bool level = rest == "sustain" || rest == "start";
        env.stages[size_t(at - std::begin(kStages))] = real(value, 0, 100) / (level ? 100 : 1);
    } else {
        const int shift = int(slot == 0 ? ModTarget::filterEnvDelay : ModTarget::pitchEnvDelay) - int(ModTarget::envDelay);
        return bindingOpcode(r, "ampeg_" + std::string(rest), value, shift, env.velocity);
    }
    return true;
}

// Filter slot opcodes (src/sfz/instrument.cpp:313-356): fil1_/fil2_ and the
// numbered cutoff/resonance spellings select the slot.
bool filterOpcode(Region& r, const std::string& op, const std::string& value)
{
    std::string name = op;
    size_t slot = 0;
    if (startsWith(name, "fil2_") || startsWith(name, "fil1_")) {
        slot = name[3] == '2';
        name.erase(3, 1);
    } else if (name == "cutoff2" || name == "resonance2" || name == "cutoff1" || name == "resonance1") {
        slot = name.back() == '2';
        name.pop_back();
    }
    static const std::map<std::string, FilterType> kTypes{
        {"lpf_1p", FilterType::low1}, {"hpf_1p", FilterType::high1}, {"lpf_2p", FilterType::low2},
        {"hpf_2p", FilterType::high2}, {"bpf_2p", FilterType::band2}, {"brf_2p", FilterType::notch2},
        {"lpf_4p", FilterType::low4}, {"hpf_4p", FilterType::high4}, {"bpf_4p", FilterType::band4},
        {"lpf_6p", FilterType::low6}, {"hpf_6p", FilterType::high6}, {"bpf_6p", FilterType::band6}};
    FilterSettings& f = r.filters[slot];
    if (name == "cutoff")
        f.cutoff = real(value, 0, 20000);
    else if (name == "resonance")
        f.resonance = real(value, -60, 96);
    else if (name == "fil_keytrack")
        f.keyTrack = real(value, -12000, 12000);
    else if (name == "fil_keycenter")
        f.keyCenter = midiKey(value);
    else if (name == "fil_veltrack")
        f.velocityTrack = real(value, -12000, 12000);
    else if (name == "fil_type") {
        const auto type = kTypes.find(value);
        if (type == kTypes.end())
            bad("unsupported filter type " + value);
        f.type = type->second;
    } else
        return false;
    if (f.type == FilterType::none)
        f.type = FilterType::low2;
    return true;
}

bool eqOpcode(Region& r, const std::string& op, const std::string& value)
{
    if (op.size() < 5 || !startsWith(op, "eq") || op[2] < '1' || op[2] > '3' || op[3] != '_')
        return false;
    EqSettings& band = r.eq[size_t(op[2] - '1')];
    const std::string_view field = std::string_view(op).substr(4);
    if (field == "freq")
        band.frequency = real(value, 0, 20000);
    else if (field == "gain")
        band.gain = real(value, -96, 96);
    else if (field == "bw")
        band.bandwidth = real(value, .001f, 12);
    else
        return false;
    band.used = true;
    return true;
}

// "varNN_..." opcodes sfizz 1.2.3 ignores.
bool ignoredVariable(const std::string& op)
{
    if (!startsWith(op, "var"))
        return false;
    const auto underscore = op.find_first_not_of("0123456789", 3);
    if (underscore == 3 || underscore == std::string::npos || op[underscore] != '_')
        return false;
    const std::string_view what = std::string_view(op).substr(underscore + 1);
    return what == "cutoff" || what == "mod" || startsWith(what, "oncc") || startsWith(what, "curvecc") ||
           startsWith(what, "eq");
}

// The region-level settings, collected from the scopes before the sample is resolved.
struct Draft {
    Region region;
    std::string sample;
    std::map<int, CcWindow> conditions;
    bool ccTriggerRange = false;
};

using Setter = std::function<void(Draft&, const std::string&)>;

const std::map<std::string, Setter>& setters()
{
    static const std::map<std::string, Setter> table = [] {
        std::map<std::string, Setter> t;
        const auto ignore = [](Draft&, const std::string&) {};
        t["sample"] = [](Draft& d, const std::string& v) { d.sample = v; };
        for (const char* label : {"key", "group_label", "master_label", "global_label", "region_label", "sw_label",
                                  "ampeg_attack_shape", "ampeg_release_shape"})
            t[label] = ignore;
        t["lokey"] = [](Draft& d, const std::string& v) { d.region.keyLow = v == "-1" ? 0 : midiKey(v); };
        t["hikey"] = [](Draft& d, const std::string& v) {
            d.region.onNotes = v != "-1";
            d.region.keyHigh = d.region.onNotes ? midiKey(v) : 127;
        };
        t["pitch_keycenter"] = [](Draft& d, const std::string& v) { d.region.keyRoot = midiKey(v); };
        t["lovel"] = [](Draft& d, const std::string& v) { d.region.velocityLow = whole(v, 0, 127); };
        t["hivel"] = [](Draft& d, const std::string& v) { d.region.velocityHigh = whole(v, 0, 127); };
        t["lochan"] = t["hichan"] = [](Draft&, const std::string& v) { (void)whole(v, 1, 16); };
        t["volume"] = [](Draft& d, const std::string& v) { d.region.volumeDb += real(v, -144, 60); };
        t["amplitude"] = [](Draft& d, const std::string& v) { d.region.amplitude *= real(v, 0, 10000) / 100; };
        t["amp_veltrack"] = [](Draft& d, const std::string& v) { d.region.velocityTrack = real(v, -100, 100) / 100; };
        t["pan"] = [](Draft& d, const std::string& v) { d.region.pan = real(v, -100, 100) / 100; };
        t["width"] = [](Draft& d, const std::string& v) { d.region.width = real(v, -100, 100) / 100; };
        t["position"] = [](Draft& d, const std::string& v) { d.region.position = real(v, -100, 100) / 100; };
        t["tune"] = t["pitch"] = [](Draft& d, const std::string& v) { d.region.tune = real(v, -1200, 1200); };
        t["transpose"] = [](Draft& d, const std::string& v) { d.region.transpose = whole(v, -127, 127); };
        t["pitch_keytrack"] = [](Draft& d, const std::string& v) { d.region.keyTrack = real(v, -1200, 1200); };
        t["bend_up"] = t["bendup"] = [](Draft& d, const std::string& v) { d.region.bendUp = real(v, -12000, 12000); };
        t["bend_down"] = t["benddown"] = [](Draft& d, const std::string& v) {
            d.region.bendDown = real(v, -12000, 12000);
        };
        t["bend_smooth"] = [](Draft& d, const std::string& v) { d.region.bendSmoothMs = whole(v, 0, 10000); };
        t["delay"] = [](Draft& d, const std::string& v) { d.region.delay = real(v, 0, 100); };
        t["oscillator_phase"] = [](Draft& d, const std::string& v) {
            d.region.oscillatorPhase = real(v, 0, 1000);  // negative (random) phases are not modelled
        };
        t["offset"] = [](Draft& d, const std::string& v) { d.region.offset = whole(v, 0, INT32_MAX - 1); };
        t["end"] = [](Draft& d, const std::string& v) { d.region.end = whole(v, 0, INT32_MAX - 1); };
        t["loop_start"] = t["loopstart"] = [](Draft& d, const std::string& v) {
            d.region.loopStart = whole(v, 0, INT32_MAX - 1);
        };
        t["loop_end"] = t["loopend"] = [](Draft& d, const std::string& v) {
            d.region.loopEnd = whole(v, 0, INT32_MAX - 1);
        };
        t["loop_crossfade"] = [](Draft& d, const std::string& v) {
            d.region.loopCrossfade = std::max(.001f, real(v, 0, 100));
        };
        t["loop_type"] = t["looptype"] = [](Draft&, const std::string& v) {
            if (v != "forward" && v != "alternate" && v != "backward")
                bad("unknown loop direction hint " + v);  // hints without effect in sfizz 1.2.3
        };
        t["loop_mode"] = [](Draft& d, const std::string& v) {
            static const std::map<std::string, LoopMode> modes{{"no_loop", LoopMode::none},
                                                              {"one_shot", LoopMode::oneShot},
                                                              {"loop_continuous", LoopMode::continuous},
                                                              {"loop_sustain", LoopMode::sustain}};
            const auto m = modes.find(v);
            if (m == modes.end())
                bad("unsupported loop mode " + v);
            d.region.loop = m->second;
        };
        t["trigger"] = [](Draft&, const std::string& v) {
            if (v != "attack")
                bad("unsupported trigger " + v);
        };
        static constexpr std::pair<const char*, EnvelopeStage> kAmp[] = {
            {"ampeg_delay", stageDelay}, {"ampeg_attack", stageAttack}, {"ampeg_hold", stageHold},
            {"ampeg_decay", stageDecay}, {"ampeg_sustain", stageSustain}, {"ampeg_start", stageStart},
            {"ampeg_release", stageRelease}};
        for (const auto& [op, stage] : kAmp) {
            const EnvelopeStage s = stage;
            t[op] = [s](Draft& d, const std::string& v) {
                d.region.amp[s] = real(v, 0, 100) / (s == stageSustain || s == stageStart ? 100 : 1);
            };
        }
        t["sustain_sw"] = [](Draft& d, const std::string& v) { d.region.sustainSwitch = onOff(v); };
        t["sustain_cc"] = [](Draft& d, const std::string& v) { d.region.sustainCC = whole(v, 0, 127); };
        t["sustain_lo"] = [](Draft& d, const std::string& v) { d.region.sustainLow = whole(v, 0, 127); };
        return t;
    }();
    return table;
}

void applyOpcode(Draft& d, const std::string& op, const std::string& value)
{
    const auto& table = setters();
    if (const auto it = table.find(op); it != table.end()) {
        it->second(d, value);
        return;
    }
    Region& r = d.region;
    if (startsWith(op, "amp_velcurve_")) {
        const int velocity = whole(op.substr(13), 0, 127);
        const float gain = real(value, 0, 1);
        auto point = std::find_if(r.velocityCurve.begin(), r.velocityCurve.end(),
                                  [&](const auto& p) { return p.first == velocity; });
        if (point == r.velocityCurve.end())
            r.velocityCurve.emplace_back(velocity, gain);
        else
            point->second = gain;
        return;
    }
    if (bindingOpcode(r, op, value, 0, r.ampVelocity) || filterOpcode(r, op, value) || eqOpcode(r, op, value) ||
        modEnvelopeOpcode(r, op, value) || ignoredVariable(op))
        return;
    if (startsWith(op, "locc") || startsWith(op, "hicc")) {
        const int cc = whole(op.substr(4), 0, 127);
        CcWindow& w = d.conditions.emplace(cc, CcWindow{cc, 0, 127}).first->second;
        (op[0] == 'l' ? w.low : w.high) = whole(value, 0, 127);
        return;
    }
    bad("unsupported SFZ opcode");
}

// Header and opcode events of an SFZ text.
class SfzReader {
public:
    using OnHeader = std::function<void(const std::string&)>;
    using OnOpcode = std::function<void(const std::string&, const std::string&)>;

    SfzReader(OnHeader header, OnOpcode opcode) : header_(std::move(header)), opcode_(std::move(opcode)) {}

    void read(const std::string& path, int depth = 0)
    {
        if (depth >= 32)
            bad("too many nested SFZ includes");
        std::ifstream in(path, std::ios::binary);
        if (!in)
            bad("cannot open SFZ " + path);
        const std::string text{std::istreambuf_iterator<char>(in), {}};
        size_t at = startsWith(text, "\xef\xbb\xbf") ? 3 : 0;
        try {
            while ((at = text.find_first_not_of(" \t\r\n", at)) != std::string::npos)
                at = next(text, at, path, depth);
        } catch (const std::exception& e) {
            const auto line = 1 + std::count(text.begin(), text.begin() + std::ptrdiff_t(std::min(at, text.size())), '\n');
            bad(path + ":" + std::to_string(line) + ": " + e.what());
        }
    }

private:
    // One element at `at`; returns where the next one may start.
    size_t next(const std::string& text, size_t at, const std::string& path, int depth)
    {
        if (text.compare(at, 2, "//") == 0)
            return text.find('\n', at);
        if (text[at] == '<') {
            const auto close = text.find('>', at + 1);
            if (close == std::string::npos)
                bad("unclosed SFZ header");
            header_(trimmed(std::string_view(text).substr(at + 1, close - at - 1)));
            return close + 1;
        }
        if (text[at] == '#') {
            const auto lineEnd = text.find('\n', at);
            directive(std::string_view(text).substr(at, lineEnd - at), path, depth);
            return lineEnd;
        }
        size_t nameEnd = at;
        while (nameEnd < text.size() && (isNameChar(text[nameEnd]) || text[nameEnd] == '$'))
            ++nameEnd;
        const auto equals = text.find_first_not_of(" \t", nameEnd);
        if (nameEnd == at || equals == std::string::npos || text[equals] != '=')
            bad("expected opcode=value");
        const std::string name = expand(text.substr(at, nameEnd - at));
        const size_t valueEnd = endOfValue(text, equals + 1);
        if (valueEnd == std::string::npos)
            bad("unclosed quoted value for " + name);
        opcode_(name, withoutQuotes(expand(trimmed(std::string_view(text).substr(equals + 1, valueEnd - equals - 1)))));
        return valueEnd;
    }

    // A value ends at a line end, a header, a comment, or whitespace followed
    // by the next `name=`; quotes protect all of these. npos: unclosed quote.
    static size_t endOfValue(const std::string& text, size_t at)
    {
        bool quoted = false;
        for (; at < text.size(); ++at) {
            const char c = text[at];
            if (c == '"')
                quoted = !quoted;
            if (quoted)
                continue;
            if (c == '\r' || c == '\n' || c == '<' || text.compare(at, 2, "//") == 0)
                break;
            if (c == ' ' || c == '\t') {
                const auto word = text.find_first_not_of(" \t", at);
                auto wordEnd = word;
                while (wordEnd < text.size() && (isNameChar(text[wordEnd]) || text[wordEnd] == '$'))
                    ++wordEnd;
                const auto after = text.find_first_not_of(" \t", wordEnd);
                if (wordEnd != word && after < text.size() && text[after] == '=')
                    break;
            }
        }
        return quoted ? std::string::npos : at;
    }

    void directive(std::string_view line, const std::string& path, int depth)
    {
        if (startsWith(line, "#include")) {
            const std::string file = trimmed(line.substr(8));
            if (file.size() < 2 || file.front() != '"' || file.back() != '"')
                bad("expected quoted #include path");
            read(joinPath(directoryOf(path), slashed(withoutQuotes(expand(file)))), depth + 1);
        } else if (startsWith(line, "#define")) {
            const std::string body = trimmed(line.substr(7));
            const auto space = body.find_first_of(" \t\r\n");
            if (body.empty() || body.front() != '$' || space == std::string::npos)
                bad("invalid #define");
            macros_[body.substr(0, space)] = expand(trimmed(std::string_view(body).substr(space)));
        } else {
            bad("unsupported SFZ directive " + std::string(line));
        }
    }

    std::string expand(std::string s) const
    {
        for (size_t at = 0; (at = s.find('$', at)) != std::string::npos;) {
            size_t end = at + 1;
            while (end < s.size() && isNameChar(s[end]))
                ++end;
            const auto macro = macros_.find(s.substr(at, end - at));
            if (macro == macros_.end())
                bad("undefined SFZ macro " + s.substr(at, end - at));
            s.replace(at, end - at, macro->second);
            at += macro->second.size();
        }
        return s;
    }

    OnHeader header_;
    OnOpcode opcode_;
    std::map<std::string, std::string> macros_;
};

class RegionBuilder {
public:
    RegionBuilder(std::string root, AudioBank& bank) : root_(std::move(root)), bank_(bank)
    {
        instrument_.ccDefault[7] = 100;
        instrument_.ccDefault[10] = 63.5f;
        instrument_.ccDefault[11] = 127;
    }

    void header(const std::string& name)
    {
        close();
        static const std::map<std::string, int> kLevels{{"global", 0}, {"master", 1}, {"group", 2}, {"region", 3},
                                                        {"control", -1}};
        const auto level = kLevels.find(name);
        if (level == kLevels.end())
            bad("unsupported SFZ header <" + name + ">");
        open_ = level->second;
        if (open_ >= 0 && open_ < 3)
            for (int s = open_; s < 4; ++s)
                scopes_[size_t(s)].clear();  // a new global, master or group starts afresh below it
        inControl_ = name == "control";
    }

    void opcode(const std::string& name, const std::string& value)
    {
        if (value.empty())
            bad("empty value for " + name);
        if (inControl_)
            control(name, value);
        else if (open_ >= 0)
            scopes_[size_t(open_)][name] = value;
        else
            bad("opcode outside SFZ header " + name);
    }

    Instrument finish(const std::string& path)
    {
        close();
        if (instrument_.regions.empty())
            bad("no playable SFZ regions in " + path);
        return std::move(instrument_);
    }

private:
    void control(const std::string& name, const std::string& value)
    {
        const auto setDefault = [&](int cc, float v) {
            if (cc < 128)
                instrument_.ccDefault[size_t(cc)] = v;
            else
                instrument_.extendedDefaults[cc] = v;
        };
        if (name == "default_path")
            defaultPath_ = slashed(value);
        else if (startsWith(name, "set_cc"))
            setDefault(whole(name.substr(6), 0, 511), real(value, 0, 127));
        else if (startsWith(name, "set_hdcc"))
            setDefault(whole(name.substr(8), 0, 511), 127 * real(value, 0, 1));
        else if (startsWith(name, "label_cc"))
            (void)whole(name.substr(8), 0, 511);
        else if (startsWith(name, "label_key"))
            (void)whole(name.substr(9), 0, 127);
        else if (name != "image" && name != "image_controls")
            bad("unsupported control opcode " + name);
    }

    // The open <region>, if any, becomes a Region.
    void close()
    {
        if (open_ != 3)
            return;
        Draft d;
        for (const Scope& scope : scopes_) {
            // `key` sets the key range and root; explicit fields of the same scope win.
            if (const auto key = scope.find("key"); key != scope.end()) {
                d.region.onNotes = key->second != "-1";
                d.region.keyLow = d.region.keyHigh = d.region.keyRoot = d.region.onNotes ? midiKey(key->second) : 60;
            }
            for (const auto& [op, value] : scope) {
                try {
                    applyOpcode(d, op, value);
                } catch (const std::exception& e) {
                    bad(op + "=" + value + ": " + e.what());
                }
            }
        }
        const bool bare = scopes_[3].empty();
        scopes_[3].clear();
        finishRegion(d, bare);
    }

    void finishRegion(Draft& d, bool bare)
    {
        Region& r = d.region;
        if (d.sample.empty() && bare)
            d.sample = "*silence";
        static const std::map<std::string, Generator> kGenerators{
            {"*sine", Generator::sine}, {"*silence", Generator::silence}, {"*noise", Generator::noise}};
        if (const auto g = kGenerators.find(d.sample); g != kGenerators.end())
            r.generator = g->second;
        else if (d.sample.empty() || d.sample.front() == '*')
            bad("missing or unsupported region sample " + d.sample);
        if (r.keyLow > r.keyHigh || r.velocityLow > r.velocityHigh)
            bad("invalid region range for " + d.sample);
        if (!r.velocityCurve.empty()) {
            std::sort(r.velocityCurve.begin(), r.velocityCurve.end());
            if (r.velocityCurve.front().first != 0)
                r.velocityCurve.insert(r.velocityCurve.begin(), {0, 0.f});
            if (r.velocityCurve.back().first != 127)
                r.velocityCurve.emplace_back(127, 1.f);
        }
        if (r.generator == Generator::none)
            resolveSample(r, d.sample);
        for (const auto& [cc, window] : d.conditions) {
            if (window.low > window.high)
                bad("invalid CC range for " + d.sample);
            r.conditions.push_back(window);
            instrument_.ccUsed[size_t(cc)] = true;
        }
        for (const auto& b : r.bindings)
            if (b.source < 128)
                instrument_.ccUsed[size_t(b.source)] = true;
        if (r.sustainSwitch)
            instrument_.ccUsed[size_t(r.sustainCC)] = true;
        const int index = int(instrument_.regions.size());
        if (r.onNotes)
            for (int key = r.keyLow; key <= r.keyHigh; ++key)
                instrument_.regionsOfKey[size_t(key)].push_back(index);
        instrument_.regions.push_back(std::move(r));
    }

    // Sample length, end and loop (src/sfz/instrument.cpp:665-679).
    void resolveSample(Region& r, const std::string& name)
    {
        r.sample = bank_.describe(joinPath(joinPath(root_, defaultPath_), slashed(name)));
        const Sample& s = bank_.samples[size_t(r.sample)];
        r.end = r.end < 0 ? s.frames : std::min(r.end + 1, s.frames);
        if (r.offset >= r.end)
            bad("offset past end of sample " + name);
        if (r.loop == LoopMode::fromFile)
            r.loop = s.loopEnd > 0 || r.loopEnd >= 0 ? LoopMode::continuous : LoopMode::none;
        if (r.loop != LoopMode::continuous && r.loop != LoopMode::sustain)
            return;
        if (r.loopStart < 0)
            r.loopStart = s.loopFirst;
        r.loopEnd = r.loopEnd < 0 ? (s.loopEnd > 0 ? s.loopEnd : r.end) : std::min(r.loopEnd + 1, r.end);
        if (r.loopStart >= r.loopEnd || r.loopEnd > r.end || r.offset >= r.loopEnd)
            bad("invalid forward loop in " + name);
    }

    std::string root_, defaultPath_;
    AudioBank& bank_;
    Instrument instrument_;
    std::array<Scope, 4> scopes_;  // global, master, group, region
    int open_ = -2;                // scope receiving opcodes; -1 control, -2 none yet
    bool inControl_ = false;
};

} // namespace

Instrument loadInstrument(const std::string& path, AudioBank& bank)
{
    RegionBuilder builder(directoryOf(slashed(path)), bank);
    SfzReader reader([&](const std::string& name) { builder.header(name); },
                     [&](const std::string& name, const std::string& value) { builder.opcode(name, value); });
    reader.read(path);
    return builder.finish(path);
}

float curveValue(int curve, float value)
{
    value = std::clamp(value, 0.f, 127.f);
    const int low = int(value);
    const int high = std::min(127, low + 1);
    const auto shape = [curve](int key) {
        const float x = key / 127.f;
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
    const float a = shape(low);
    return a + (value - low) * (shape(high) - a);
}

} // namespace clay
