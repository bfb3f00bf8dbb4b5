// Sensor and ISP parameters for the MI355X renderer, with the names, defaults and provenance of the reference's
// gpu/sensor.js (sensorParams 17-84, describe 87-92, analogGain 94-97, AutoExposure 127-155, rawNoiseDN 180-182,
// updateAWB 185-190). Where the reference packs a 256-B WebGPU uniform (packSensorUniform 158-177), this module
// builds the two parameter objects the addon passes to vt_sensor and vt_isp; the numbers derived on the way (optics
// constant, ESF remainder, maximum DN, glare sigma at the pyramid level, mode codes) are the same.

const PLANCK = 6.62607015e-34, LIGHT = 299792458;
const CHANNEL_NM = [600, 540, 460];                 // wavelengths standing for the R, G, B channels
export const CFA_CODES = Object.freeze({ RGGB: 0, GRBG: 1, GBRG: 2, BGGR: 3 });
export const ANALOG_GAIN_MAX = 8;                   // UVC gain 0..100 onto 1..8x (assumed, as the reference)

const isBox = (x) => x && typeof x === 'object' && !Array.isArray(x) && 'value' in x;
const valueOf = (x) => (isBox(x) ? x.value : x);
const provOf = (x) => (isBox(x) ? { status: x.status, source: x.source } : { status: 'default', source: 'js/sensor.mjs' });

/** Sensor + ISP parameters from physics.yaml's `sensor` and `isp` blocks ({value, status, source} leaves or plain
 * numbers), with provenance; `over` overrides any field of the result (gpu/sensor.js:17-84). */
export function sensorParams(physics = {}, over = {}) {
  const s = physics.sensor || {}, isp = physics.isp || {}, lens = s.lens || {};
  const uvc = s.uvc || {}, ae = s.auto_exposure || {}, ri = lens.relative_illumination || {};
  const provenance = {};
  const pick = (key, x, dflt) => {
    provenance[key] = provOf(x);
    const v = valueOf(x);
    return v === undefined || v === null ? dflt : v;
  };
  const pitch = pick('pixel_pitch', s.pixel_pitch, 3e-6);
  const qe = pick('qe_peak', s.qe_peak, [0.45, 0.55, 0.5]);
  const mBox = s.response_matrix ? s.response_matrix.M : undefined;
  let M = pick('response_matrix', mBox, null);
  const tagged = /placeholder/i.test(String((mBox && mBox.source) || ''));
  if (!M || tagged || Math.max(...M.flat().map(Math.abs)) < 1e3) {
    // e-/s per W m^-2 on the pixel: area x QE x photons per joule at the channel's wavelength
    const g = CHANNEL_NM.map((nm, c) => (pitch * pitch * qe[c] * nm * 1e-9) / (PLANCK * LIGHT));
    M = [[g[0], 0, 0], [0, g[1], 0], [0, 0, g[2]]];
    provenance.response_matrix = { status: 'derived', source: 'js/sensor.mjs: pixel_pitch^2 x qe_peak x lambda/(h c)' };
  }
  const gainRange = pick('analog_gain_range', s.analog_gain_range, [1, ANALOG_GAIN_MAX]) || [1, ANALOG_GAIN_MAX];
  const p = {
    responseMatrix: M,
    fNumber: pick('f_number', lens.f_number, 2.0),
    transmission: pick('transmission', lens.transmission, 0.9),
    pupilExponent: pick('pupil_obliquity_exponent', ri.pupil_obliquity_exponent, 1.0),
    vignettingThetaDeg: pick('vignetting_theta_deg', ri.vignetting_theta_deg, [0, 90]),
    vignetting: pick('vignetting', ri.vignetting, [1, 1]),
    esfSigmaPx: pick('esf_sigma_px', lens.esf_sigma_px, 0.7),
    esfIncludesAperture: true,
    readNoise: pick('read_noise', s.read_noise, 3.0),
    fullWell: pick('full_well', s.full_well, 10000),
    darkCurrent: pick('dark_current', s.dark_current, 0),
    conversionGain: pick('conversion_gain_base', s.conversion_gain_base, 0.33),
    adcBits: pick('adc_bits', s.adc_bits, 12),
    blackLevel: pick('black_level', s.black_level, 0),
    cfa: pick('bayer_order', s.bayer_order, 'GRBG'),
    exposureUnit: pick('exposure_unit', uvc.exposure_unit, 1e-4),
    exposureRange: pick('exposure_range', uvc.exposure_range, [1, 5000]),
    gainRange: pick('gain_range', uvc.gain_range, [0, 100]),
    fps: pick('frame_rate', uvc.frame_rate, 20),
    aeTarget: pick('target_mean', ae.target_mean, 0.45),
    aeDamping: pick('damping', ae.damping, 0.3),
    wb: pick('white_balance_gains', isp.white_balance_gains, [1, 1, 1]),
    ccm: pick('ccm', isp.ccm, [[1, 0, 0], [0, 1, 0], [0, 0, 1]]),
    tone: pick('tone_curve', isp.tone_curve, 'srgb'),
    sCurveC: pick('s_curve_c', isp.s_curve_c, null),
    blackOffset: pick('black_offset_linear', isp.black_offset_linear, 0),
    demosaic: pick('demosaic', isp.demosaic, 'bilinear'),
    noiseReduction: pick('noise_reduction', isp.noise_reduction, 'none'),
    nrSigmaPx: pick('noise_reduction_sigma_px', isp.noise_reduction_sigma_px, 0.4),
    nrRangeScale: 2.0,
    glareFraction: pick('veiling_glare.fraction', lens.veiling_glare && lens.veiling_glare.fraction, 0),
    glareSigmaFrac: pick('veiling_glare.sigma_frac_diag', lens.veiling_glare && lens.veiling_glare.sigma_frac_diag, 0.25),
    glareLevel: 16,
    gamma: pick('gamma', isp.gamma, 2.2),
    sharpenAmount: pick('sharpening.amount', isp.sharpening && isp.sharpening.amount, 0),
    sharpenRadius: pick('sharpening.radius_px', isp.sharpening && isp.sharpening.radius_px, 1),
    chroma: pick('chroma_subsampling', isp.chroma_subsampling, '4:2:0'),
    aeMaxStep: pick('max_step', ae.max_step, 2.0),
    awb: !!pick('white_balance_auto_default', uvc.white_balance_auto_default, true),
    awbDamping: 0.3,
    noise: true,
    analogGainMax: gainRange[1],
    ...over,
  };
  provenance.nrRangeScale = { status: 'assumed', source: 'bilateral range sigma = 2 x the modelled noise' };
  provenance.glareLevel = { status: 'assumed', source: 'glare on a 1/16 pyramid level' };
  provenance.gainMapping = { status: 'assumed', source: 'UVC gain 0..100 -> analog_gain_range, linear' };
  provenance.awbMethod = { status: 'assumed', source: 'grey-world AWB on the raw channel means, damped 0.3 per frame' };
  const e0 = valueOf(uvc.exposure_default);
  p.exposureUnits0 = e0 === undefined || e0 === null ? 10 : e0;
  p.provenance = provenance;
  return p;
}

/** Counts of the parameters by status, and the provenance table (gpu/sensor.js:87-92). */
export function describe(p) {
  const counts = {};
  for (const v of Object.values(p.provenance)) {
    const k = v.status || 'unknown';
    counts[k] = (counts[k] || 0) + 1;
  }
  return { counts, provenance: p.provenance };
}

/** UVC gain units -> analogue gain (linear map of gainRange onto 1..analogGainMax). */
export function analogGain(p, uvcGain) {
  const [g0, g1] = p.gainRange;
  const u = Math.min(g1, Math.max(g0, uvcGain));
  return 1 + ((u - g0) / Math.max(1, g1 - g0)) * (p.analogGainMax - 1);
}

/** UVC-style auto exposure: whole exposure units first (up to the frame time), then analogue gain
 * (gpu/sensor.js:127-155). update(meanLuma) takes the last frame's mean output luma in 0..1. */
export class AutoExposure {
  constructor(p, { exposureUnits = 10, gainUnits = 0, enabled = true } = {}) {
    this.p = p;
    this.units = exposureUnits;
    this.gainUnits = gainUnits;
    this.enabled = enabled;
    this.product = this.units * p.exposureUnit * analogGain(p, gainUnits);
  }
  get exposure() { return this.units * this.p.exposureUnit; }
  get gain() { return analogGain(this.p, this.gainUnits); }
  update(meanLuma) {
    if (!this.enabled) return { exposure: this.exposure, gain: this.gain };
    const p = this.p, unit = p.exposureUnit, maxStep = p.aeMaxStep || 2;
    const gamma = p.tone === 'linear' ? 1 : 2.2;
    const ratio = Math.min(8, Math.max(1 / 8, meanLuma > 1e-4 ? Math.pow(p.aeTarget / meanLuma, gamma) : 8));
    this.product *= Math.min(maxStep, Math.max(1 / maxStep, Math.pow(ratio, p.aeDamping)));
    const maxUnits = Math.min(p.exposureRange[1], Math.floor(1 / (p.fps * unit)));
    this.units = Math.max(p.exposureRange[0], Math.min(maxUnits, Math.round(this.product / unit)));
    const g = Math.max(1, Math.min(p.analogGainMax, this.product / (this.units * unit)));
    const [g0, g1] = p.gainRange;
    this.gainUnits = Math.round(g0 + ((g - 1) / (p.analogGainMax - 1)) * (g1 - g0));
    this.product = Math.max(p.exposureRange[0] * unit, Math.min(maxUnits * unit * p.analogGainMax, this.product));
    return { exposure: this.exposure, gain: this.gain };
  }
}

/** Modelled noise std in DN of one raw pixel at a mean of `e` electrons: shot + read, times gain. */
export function rawNoiseDN(p, e, gain) {
  return Math.sqrt(Math.max(0, e) + p.readNoise * p.readNoise) * p.conversionGain * gain;
}

/** Grey-world white balance: gains that equalise the raw channel means, moved by `damping` in log space. */
export function updateAWB(prev, meanRGB, damping = 0.3) {
  const g = prev || [1, 1, 1];
  if (!(meanRGB[0] > 1e-4 && meanRGB[1] > 1e-4 && meanRGB[2] > 1e-4)) return g;
  const target = [meanRGB[1] / meanRGB[0], 1, meanRGB[1] / meanRGB[2]];
  return g.map((x, c) => Math.exp(Math.log(x) + damping * (Math.log(target[c]) - Math.log(x))));
}

/** The raw stage's parameters for the addon (vt_sensor_params). */
export function nativeSensor(p, { width, height, frame = 0, seed = 0, exposure, gain }) {
  const sOpt = p.esfIncludesAperture ? Math.sqrt(Math.max(0, p.esfSigmaPx * p.esfSigmaPx - 1 / 12)) : p.esfSigmaPx;
  const F = p.glareLevel || 16;
  const code = CFA_CODES[p.cfa];
  return {
    cfa: code === undefined ? 1 : code, frame: frame >>> 0, seed: seed >>> 0,
    opticsK: (Math.PI * p.transmission) / (4 * p.fNumber * p.fNumber), esfSigma: sOpt, exposure, gain,
    m: p.responseMatrix.flat(), readNoise: p.readNoise, fullWell: p.fullWell, dark: p.darkCurrent,
    convGain: p.conversionGain, maxDn: Math.pow(2, p.adcBits) - 1, black: p.blackLevel, noise: p.noise ? 1 : 0,
    pupilExp: p.pupilExponent, glareFraction: p.glareFraction || 0,
    glareSigma: (p.glareSigmaFrac * Math.hypot(width, height)) / F, glareLevel: F,
  };
}

/** The ISP's parameters for the addon (vt_isp_params): the reference's mode strings as codes. */
export function nativeIsp(p) {
  const tone = p.tone === 'srgb' ? (p.sCurveC ? 3 : 0) : p.tone === 'linear' ? 2 : 1;
  return {
    demosaic: /hamilton/i.test(String(p.demosaic)) ? 2 : /edge/.test(String(p.demosaic)) ? 1 : 0,
    nr: /edge|bilateral|on/.test(String(p.noiseReduction)) ? 1 : 0,
    tone, chroma420: String(p.chroma).startsWith('4:2:0') ? 1 : 0,
    wb: Array.from(p.wb), ccm: p.ccm.flat(), gamma: p.gamma, sCurveC: p.sCurveC || 1, blackOffset: p.blackOffset || 0,
    nrSigmaPx: p.nrSigmaPx || 0, nrRangeScale: p.nrRangeScale || 2, sharpenAmount: p.sharpenAmount,
    sharpenRadius: p.sharpenRadius,
  };
}
