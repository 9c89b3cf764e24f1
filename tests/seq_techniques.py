"""Non-Cartesian and echo-train variants of tests/golden/psd_gre3d_small.xml
for the sequence-compiler tests: the GRE fixture with its phase-encode,
readout and ADC elements swapped for the EPI, FSE, radial or spiral ones
(Macro/SeqElem/{GyPE,GxR,ADC}/*.m) and a VCtl derived by
mrilab_amd.prescan as the scan set-up derives it (DoPreScan.m:33-109)."""
import os

from mrilab_amd import prescan

HERE = os.path.dirname(os.path.abspath(__file__))
PSD = open(os.path.join(HERE, "golden", "psd_gre3d_small.xml")).read()
ON, OFF = "$1'on','off'", "$2'on','off'"

BASE = dict(TR=10e-3, TE=4e-3, FlipAng=20.0, FOVFreq=0.064, FOVPhase=0.064, SliceThick=0.016, ResFreq=16.0,
            ResPhase=16.0, SliceNum=2.0, BandWidth=100e3, MaxGrad=0.05, MaxSlewRate=200.0, MinUpdRate=1e-6,
            TEAnchor="Middle", MultiTransmit="off")


def _swap(psd, tag, new):
    i = psd.index("<" + tag)
    return psd[:i] + new + psd[psd.index("/>", i) + 2:]


def epi(shifting="off", etl=8.0, shots=2.0):
    psd = _swap(PSD, "GyCartesian", f'<GyEPI Gy1Sign="-1" Gy2Sign="1" Notes="blips" Switch="{ON}" t1Start="CV1" '
                                     f't2Middle="VCtl.TE"/>')
    psd = _swap(psd, "GxCartesian", f'<GxEPI Gx1Sign="-1" Gx2Sign="1" Notes="zigzag" Switch="{ON}" t1Start="CV1" '
                                    f't2Middle="VCtl.TE"/>')
    psd = _swap(psd, "ADCCartesian", f'<ADCEPI Notes="lobes" Switch="{ON}" tMiddle="VCtl.TE"/>')
    psd = psd.replace('Gz2Sign="-1"', 'Gz2Sign="0"')  # no slice rewinder inside the echo train
    v = prescan.derive(dict(BASE, ResFreq=16.0, EPI_ETL=etl, EPI_ESP=0.4e-3, EPI_ShotNum=shots,
                            EPI_EchoShifting=shifting))
    return psd, v


def fse(etl=4.0, shots=2.0):
    rf180 = (f'<rfSinc AnchorTE="{OFF}" Apod="$1\'Non\',\'Hamming\',\'Hanning\'" CoilID="1" '
             f'DupSpacing="VCtl.FSE_ESP" Duplicates="VCtl.FSE_ETL" FA="180" Notes="refocusing" Switch="{ON}" '
             f'TBP="4" dt="10e-6" rfFreq="0" rfPhase="pi/2" tEnd="VCtl.FSE_ESP/2+0.4e-3" tStart="VCtl.FSE_ESP/2"/>')
    psd = PSD.replace("</rf>", rf180 + "</rf>", 1)
    # one slice: the slice encodes have zero area (the line may not be empty)
    psd = _swap(psd, "GyCartesian", f'<GyFSE Gy1Sign="1" Gy2Sign="-1" Notes="encode and rewind" Switch="{ON}" '
                                    f'tGy1="0.5e-3" tGy2="0.5e-3" tMiddle="VCtl.TE" tOffset="1.5e-3"/>')
    psd = _swap(psd, "GxCartesian", f'<GxFSE Gx1Sign="1" Gx2Sign="1" Notes="readouts" Switch="{ON}" '
                                    f't1Start="CV1" t2Middle="VCtl.TE"/>')
    psd = _swap(psd, "ADCCartesian", f'<ADCFSE Notes="echoes" Switch="{ON}" tMiddle="VCtl.TE"/>')
    rev = (f'<ExtBit DupSpacing="VCtl.FSE_ESP" Duplicates="VCtl.FSE_ETL" Ext="2" Notes="reverse k" '
           f'Switch="{ON}" tStart="VCtl.FSE_ESP/2+0.2e-3"/>')
    psd = psd.replace("</Ext>", rev + "</Ext>", 1)
    esp = 4e-3
    v = prescan.derive(dict(BASE, TR=20e-3, TE=(etl // 2 + 1) * esp, SliceNum=1.0, FSE_ETL=etl, FSE_ESP=esp,
                            FSE_ShotNum=shots))
    return psd, v


def radial(pattern="Linear", ang_range="pi", spokes=8.0, samples=17.0):
    def g(ax):
        return (f'<G{ax}Radial G{ax}1Sign="-1" G{ax}2Sign="1" G{ax}3Sign="0" Notes="spoke" Switch="{ON}" '
                f't1Start="CV1" t2Middle="VCtl.TE" t3Start="VCtl.TE+CV1" tRamp="100e-6"/>')
    psd = _swap(PSD, "GyCartesian", g("y"))
    psd = _swap(psd, "GxCartesian", g("x"))
    psd = _swap(psd, "ADCCartesian", f'<ADCRadial Notes="spoke samples" Switch="{ON}" tMiddle="VCtl.TE"/>')
    v = prescan.derive(dict(BASE, R_SpokeNum=spokes, R_AngPattern=pattern, R_AngRange=ang_range,
                            R_SampPerSpoke=samples))
    return psd, v


def spiral(gradient=0.01, shots=4.0, dt=4e-6):
    def g(ax):
        return f'<G{ax}Spiral Notes="interleave" Switch="{ON}" dt="{dt}" tStart="VCtl.TE"/>'
    psd = _swap(PSD, "GyCartesian", g("y"))
    psd = _swap(psd, "GxCartesian", g("x"))
    psd = _swap(psd, "ADCCartesian", f'<ADCSpiral Notes="spiral samples" Switch="{ON}" tStart="VCtl.TE"/>')
    psd = psd.replace('Gz2Sign="-1"', 'Gz2Sign="0"')
    v = prescan.derive(dict(BASE, TR=40e-3, TE=2e-3, S_Lamda=2.1, S_SlewRate=150.0, S_SlewRate0=100.0,
                            S_Gradient=gradient, S_ShotNum=shots))
    return psd, v


def rf_only(rf, gz='<GzTrapezoid DupSpacing="0" Duplicates="1" GzAmp="0" Notes="z" Switch="$1\'on\',\'off\'" '
                 'sRamp="2" tEnd="VCtl.TR-0.5e-3" tRamp="10e-6" tStart="VCtl.TR-0.6e-3"/>',
            gy='<GyAreaTrapezoid Area="0" DupSpacing="0" Duplicates="1" Notes="y" Switch="$1\'on\',\'off\'" '
               'tEnd="VCtl.TR-0.5e-3" tStart="VCtl.TR-0.6e-3"/>',
            gx='<GxAreaTrapezoid2 Area="0" DupSpacing="0" Duplicates="1" Notes="x" Switch="$1\'on\',\'off\'" '
               'tStart="VCtl.TR-0.6e-3"/>'):
    """A one-TR PSD of one rf element, zero-area gradient elements (the
    lines may not be empty) and an ADC at the end of the TR."""
    psd = (f'<?xml version="1.0" encoding="utf-8"?><MRiLabSeq Name="rf" Notes="one pulse"><CVs CV1="0"/>'
           f'<Pulses Freq="1" Notes="TR" Switch="{ON}" TREnd="Inf" TRStart="1" tE="VCtl.TR" tS="0">'
           f'<rf>{rf}</rf><GzSS>{gz}</GzSS><GyPE>{gy}</GyPE><GxR>{gx}</GxR>'
           f'<ADC><ADCCartesian DupSpacing="0" Duplicates="1" Notes="adc" Switch="{ON}" '
           f'tMiddle="VCtl.TR-0.3e-3"/></ADC>'
           f'<Ext><ExtBit DupSpacing="0" Duplicates="1" Ext="1" Notes="reset k" Switch="{ON}" tStart="0"/></Ext>'
           f'</Pulses></MRiLabSeq>')
    v = prescan.derive(dict(BASE, TR=12e-3, SliceNum=1.0, ResFreq=4.0, ResPhase=2.0))
    v["FirstPhNum"] = v["SecondPhNum"] = 1.0
    return psd, v


def hs(scale=1.0, t0=0.5e-3):
    return (f'<rfHyperbolicSecant Adiab="4.9" AnchorTE="{OFF}" CoilID="1" DupSpacing="0" Duplicates="1" '
            f'MaxB1="{14e-6 * scale}" Notes="HS" Switch="{ON}" TBP="10" dt="10e-6" rfPhase="0" tEnd="{t0 + 8e-3}" '
            f'tStart="{t0}"/>')


def tanhtan(scale=1.0):
    return (f'<rfTanhTan AnchorTE="{OFF}" CoilID="1" DupSpacing="0" Duplicates="1" MaxB1="{9e-4 * scale}" '
            f'Notes="tt" Switch="{ON}" TBP="50" dt="1e-6" rfPhase="0.2" tEnd="60e-6" tStart="10e-6"/>')


def bir(flag, dups=1):
    k = {"BIR-1": 1, "BIR-2": 2, "BIR-4": 3}[flag]
    return (f'<rfBIR AnchorTE="{OFF}" BIRFlag="${k}\'BIR-1\',\'BIR-2\',\'BIR-4\'" Beta="1.4711" CoilID="1" '
            f'DupSpacing="4e-3" Duplicates="{dups}" Lambda="10" MaxB1="10e-6" MaxFreq="10e3" Notes="bir" '
            f'Switch="{ON}" dt="10e-6" tEnd="4e-3" tStart="1e-3"/>')


def biref(flag):
    k = {"BIREF-1": 1, "BIREF-2a": 2, "BIREF-2b": 3}[flag]
    return (f'<rfBIREF AnchorTE="{OFF}" BIREFFlag="${k}\'BIREF-1\',\'BIREF-2a\',\'BIREF-2b\'" CoilID="1" '
            f'DupSpacing="0" Duplicates="1" MaxB1="10e-6" MaxFreq="10e3" Notes="biref" Switch="{ON}" dt="10e-6" '
            f'tEnd="3e-3" tStart="1e-3"/>')
