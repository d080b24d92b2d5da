// Phasor::ATA module (reference src/modules/phasor/generic.cc:11-96 and
// ata.cc:53-213).
//
// Delays: the reference computes them with radiointerferometryc99 + ERFA
// (ata.cc:105-173), neither of which exists here.  This module restates the
// geometry from first principles, in FP64 on the host:
//   1. antenna ECEF -> baseline from the array reference position (WGS84),
//      rotated by the reference longitude into the local XYZ frame
//      (X: meridian at the equator, Y: east, Z: north pole);
//   2. source (RA, Dec) of J2000 precessed to the mean equator of date
//      (IAU 1976 angles zeta, z, theta), hour angle from the Greenwich mean
//      sidereal time of UT1 (IAU 1982 expression) plus the east longitude;
//   3. w = cos(dec) cos(H) X - cos(dec) sin(H) Y + sin(dec) Z, and
//      tau = (w_ant - w_ref) / c, beam delays minus boresight delays.
// Nutation and aberration (arcsecond-level) are not modelled; see the test
// module for the tolerance this implies.  The phasors themselves are made on
// the GPU from the delays with an FP64 phase (blade_phasor_ata).
#include "blade/base.hh"

#include <cmath>
#include <string>

namespace bl {
void set_error(const std::string& msg);   // blade_amd_last_error() message (runtime.cc)
}

namespace Blade::Modules::Phasor {

#undef BL_LOG_DOMAIN
#define BL_LOG_DOMAIN "M::PHASOR"

template <typename OT>
Generic<OT>::Generic(const Config& config, const Input& input, const Stream&)
    : config(config), input(input), numberOfFrequencySteps(0), currentFrequencyStep(0) {
    if (config.numberOfAntennas != config.antennaPositions.size()) {
        BL_FATAL("Number of Antennas configuration ({}) mismatches the number of antenna positions ({}).",
                 config.numberOfAntennas, config.antennaPositions.size());
        throw Result::ERROR;
    }
    if (config.referenceAntennaIndex >= config.numberOfAntennas) {
        BL_FATAL("Reference Antenna Index ({}) is larger than the number of antennas ({}).",
                 config.referenceAntennaIndex, config.numberOfAntennas);
        throw Result::ERROR;
    }
    if (Profiler::IsCapturing()) {
        BL_WARN("Capturing: Early setup return.");
        return;
    }
    // CF16 headroom warning (reference generic.cc:52-86).
    const F64 max_value = 65500.0 / (config.numberOfAntennas * 127.0);
    F64 max_cal = 0.0, min_cal = 0.0;
    for (const auto& c : config.antennaCalibrations) {
        max_cal = std::max({max_cal, c.real(), c.imag()});
        min_cal = std::min({min_cal, c.real(), c.imag()});
    }
    if (max_value < max_cal || -max_value > min_cal) {
        BL_WARN("Overflow Possible! At least one calibration value is smaller or larger ({}, {}) "
                "than what CF16 can hold ({}, {}) with current configuration parameters.",
                min_cal, max_cal, -max_value, max_value);
    }
    BL_INFO("Type: {} -> {}", "N/A", TypeInfo<OT>::name);
    BL_INFO("Observation Frequency (Hz): {}", config.observationFrequencyHz);
    BL_INFO("Channel Bandwidth (Hz): {}", config.channelBandwidthHz);
    BL_INFO("Total Bandwidth (Hz): {}", config.totalBandwidthHz);
    BL_INFO("Frequency Start Index: {}", config.frequencyStartIndex);
    BL_INFO("Reference Antenna Index: {}", config.referenceAntennaIndex);
}

namespace {

constexpr F64 kPi = 3.14159265358979323846;
constexpr F64 kArcsec = kPi / (180.0 * 3600.0);

// WGS84 geodetic -> ECEF (metres).
XYZ lla_to_ecef(const LLA& p) {
    const F64 a = 6378137.0, f = 1.0 / 298.257223563, e2 = f * (2.0 - f);
    const F64 s = std::sin(p.LAT), c = std::cos(p.LAT);
    const F64 N = a / std::sqrt(1.0 - e2 * s * s);
    return {(N + p.ALT) * c * std::cos(p.LON), (N + p.ALT) * c * std::sin(p.LON),
            (N * (1.0 - e2) + p.ALT) * s};
}

// Greenwich mean sidereal time (radians) of a UT1 Julian date (IAU 1982).
F64 gmst(F64 jd_ut1) {
    const F64 d = jd_ut1 - 2451545.0;
    const F64 T = d / 36525.0;
    F64 sec = 67310.54841 + (876600.0 * 3600.0 + 8640184.812866) * T + 0.093104 * T * T -
              6.2e-6 * T * T * T;
    F64 r = std::fmod(sec / 86400.0 * 2.0 * kPi, 2.0 * kPi);
    return r < 0 ? r + 2.0 * kPi : r;
}

// Precess J2000 mean (ra, dec) to the mean equator and equinox of date (IAU 1976).
void precess(F64 jd, F64& ra, F64& dec) {
    const F64 T = (jd - 2451545.0) / 36525.0;
    const F64 zeta = (2306.2181 * T + 0.30188 * T * T + 0.017998 * T * T * T) * kArcsec;
    const F64 z = (2306.2181 * T + 1.09468 * T * T + 0.018203 * T * T * T) * kArcsec;
    const F64 theta = (2004.3109 * T - 0.42665 * T * T - 0.041833 * T * T * T) * kArcsec;
    const F64 A = std::cos(dec) * std::sin(ra + zeta);
    const F64 B = std::cos(theta) * std::cos(dec) * std::cos(ra + zeta) - std::sin(theta) * std::sin(dec);
    const F64 C = std::sin(theta) * std::cos(dec) * std::cos(ra + zeta) + std::cos(theta) * std::sin(dec);
    ra = std::atan2(A, B) + z;
    dec = std::asin(std::max(-1.0, std::min(1.0, C)));
}

// w of every antenna relative to the reference antenna, divided by c, for a
// J2000 (ra, dec) at a UTC Julian date: the delay (s) of each antenna's
// signal after the reference antenna's.
void position_delays(const std::vector<XYZ>& xyz, U64 ref, F64 lon, F64 ra, F64 dec, F64 jd,
                     F64 dut1, F64* delays) {
    const F64 jd_ut1 = jd + dut1 / 86400.0;
    precess(jd, ra, dec);
    const F64 H = gmst(jd_ut1) + lon - ra;
    const F64 cd = std::cos(dec), sd = std::sin(dec), ch = std::cos(H), sh = std::sin(H);
    const XYZ& r = xyz[ref];
    const F64 wr = cd * ch * r.X - cd * sh * r.Y + sd * r.Z;
    for (U64 a = 0; a < xyz.size(); a++) {
        const XYZ& p = xyz[a];
        const F64 w = cd * ch * p.X - cd * sh * p.Y + sd * p.Z;
        delays[a] = (w - wr) / BL_PHYSICAL_CONSTANT_C;
    }
}

// ECEF antenna positions -> baselines from the array reference position in
// the local XYZ frame (rotation by the reference longitude).
std::vector<XYZ> local_xyz(const XYZ* ecef, U64 A, const LLA& refPos) {
    const XYZ ref = lla_to_ecef(refPos);
    const F64 cl = std::cos(refPos.LON), sl = std::sin(refPos.LON);
    std::vector<XYZ> xyz(A);
    for (U64 i = 0; i < A; i++) {
        const XYZ& p = ecef[i];
        const F64 bx = p.X - ref.X, by = p.Y - ref.Y, bz = p.Z - ref.Z;
        xyz[i] = {bx * cl + by * sl, -bx * sl + by * cl, bz};
    }
    return xyz;
}

}  // namespace

// Beam-minus-boresight delays [B][A] (reference ata.cc:105-173).
void ataDelays(const std::vector<XYZ>& xyz, U64 ref, F64 lon, const RA_DEC& boresight,
               const RA_DEC* beams, U64 B, F64 jd, F64 dut1, F64* delays) {
    const U64 A = xyz.size();
    std::vector<F64> bore(A);
    position_delays(xyz, ref, lon, boresight.RA, boresight.DEC, jd, dut1, bore.data());
    for (U64 b = 0; b < B; b++) {
        F64* d = delays + b * A;
        position_delays(xyz, ref, lon, beams[b].RA, beams[b].DEC, jd, dut1, d);
        for (U64 a = 0; a < A; a++) d[a] -= bore[a];
    }
}

template <typename OT>
ATA<OT>::ATA(const typename Generic<OT>::Config& config, const typename Generic<OT>::Input& input,
             const Stream& stream)
    : Generic<OT>(config, input, stream) {
    const U64 calChannels = config.antennaCalibrations.shape().numberOfFrequencyChannels();
    if (calChannels != this->config.numberOfFrequencyChannels) {
        BL_FATAL("Number of frequency channels ({}) has to be equal to the antenna calibrations "
                 "frequency channels ({}).", this->config.numberOfFrequencyChannels, calChannels);
        throw Result::ERROR;
    }
    this->numberOfFrequencySteps = calChannels / this->config.numberOfFrequencyChannels;
    if (getConfigCalibrationShape() != config.antennaCalibrations.shape()) {
        BL_FATAL("Shape of antenna calibrations {} doesn't match with the expected dimensions {}.",
                 config.antennaCalibrations.shape(), getConfigCalibrationShape());
        throw Result::ERROR;
    }
    // ECEF -> array-centred XYZ (rotation by the reference longitude).
    antennasXyz = local_xyz(config.antennaPositions.data(), config.numberOfAntennas,
                            config.arrayReferencePosition);
    this->output.phasors = PhasorTensor<Device::CUDA, OT>(getOutputPhasorsShape());
    this->output.delays = DelayTensor<Device::CPU, F64>(getOutputDelaysShape());
    deviceDelays = DelayTensor<Device::CUDA, F64>(getOutputDelaysShape());
    deviceCalibrations = ArrayTensor<Device::CUDA, CF64>(config.antennaCalibrations.shape());
    if (!Profiler::IsCapturing()) {
        BL_CHECK_THROW(Copy(deviceCalibrations, config.antennaCalibrations, stream));
        BL_HIP_CHECK_THROW(hipStreamSynchronize(stream.handle));
        BL_HIP_CHECK_THROW(hipEventCreateWithFlags(&uploaded, hipEventDisableTiming));
    }
    BL_INFO("Phasors Shape: {} -> {}", "N/A", this->getOutputPhasors().shape());
    BL_INFO("Delays Shape: {} -> {}", "N/A", this->getOutputDelays().shape());
}

template <typename OT>
ATA<OT>::~ATA() {
    if (uploaded) (void)hipEventDestroy(uploaded);
}

template <typename OT>
Result ATA<OT>::process(const U64&, const Stream& stream) {
    const auto& cfg = this->config;
    const U64 A = cfg.numberOfAntennas, B = cfg.beamCoordinates.size();
    const F64 jd = this->input.blockJulianDate[0];
    const F64 dut1 = this->input.blockDut1[0];
    // The previous block's upload (possibly on the other stream) must have
    // read the pinned host delays before they are overwritten.
    if (uploaded) BL_HIP_CHECK(hipEventSynchronize(uploaded));
    ataDelays(antennasXyz, cfg.referenceAntennaIndex, cfg.arrayReferencePosition.LON,
              cfg.boresightCoordinate, cfg.beamCoordinates.data(), B, jd, dut1,
              this->output.delays.data());
    BL_CHECK(Copy(deviceDelays, this->output.delays, stream));
    if (uploaded) BL_HIP_CHECK(hipEventRecord(uploaded, stream.handle));
    BL_ABI_CHECK(blade_phasor_ata(deviceDelays.data(), deviceCalibrations.data(),
                                  this->output.phasors.data(), TypeInfo<OT>::abi, B, A,
                                  cfg.numberOfFrequencyChannels, cfg.numberOfPolarizations,
                                  cfg.observationFrequencyHz, cfg.channelBandwidthHz,
                                  cfg.totalBandwidthHz, cfg.frequencyStartIndex, stream.abi()));
    this->currentFrequencyStep = (this->currentFrequencyStep + 1) % this->numberOfFrequencySteps;
    return Result::SUCCESS;
}

template class BLADE_API Generic<CF32>;
template class BLADE_API Generic<CF64>;
template class BLADE_API ATA<CF32>;
template class BLADE_API ATA<CF64>;

}  // namespace Blade::Modules::Phasor

extern "C" int blade_ata_delays(const double* antenna_ecef, uint64_t A, uint64_t reference_antenna,
                                const double* reference_lla, const double* boresight_radec,
                                const double* beam_radec, uint64_t B, double julian_date,
                                double dut1, double* delays) {
    using namespace Blade;
    using namespace Blade::Modules::Phasor;
    if (A == 0 || B == 0) return BLADE_SUCCESS;
    if (!antenna_ecef || !reference_lla || !boresight_radec || !beam_radec || !delays) {
        bl::set_error("ata delays: null pointer");
        return BLADE_ERROR;
    }
    if (reference_antenna >= A) {
        bl::set_error("ata delays: reference antenna index is not below the number of antennas");
        return BLADE_ERROR;
    }
    const LLA ref{reference_lla[0], reference_lla[1], reference_lla[2]};
    const std::vector<XYZ> xyz = local_xyz(reinterpret_cast<const XYZ*>(antenna_ecef), A, ref);
    const RA_DEC bore{boresight_radec[0], boresight_radec[1]};
    ataDelays(xyz, reference_antenna, ref.LON, bore, reinterpret_cast<const RA_DEC*>(beam_radec), B,
              julian_date, dut1, delays);
    return BLADE_SUCCESS;
}
