"""Writes pmath_table.h, the tables of micm_amd/csrc/pmath.h:

    python micm_amd/csrc/gen_pmath_table.py > micm_amd/csrc/pmath_table.h

The functions reproduce the exp / log / pow of ARM Optimized Routines
(https://github.com/ARM-software/optimized-routines, math/exp.c, log.c, pow.c;
MIT OR Apache-2.0 WITH LLVM-exception), which the GNU C library has used since
2.28.  Their tables are defined by published formulas:

* exp: for k = 0..127, H = 2^(k/128) and T = (2^(k/128) - H) / H, each rounded to
  double; the table holds the bits of T and the bits of H minus (k << 45).
* pow's logarithm: z in [0x1.69555p-1, 0x1.69555p0) split by the top seven
  significand bits into 128 subintervals; with `center` the subinterval's midpoint,
  1/c = round(128 / center) / 128 below 1 and round(256 / center) / 256 above;
  logc = round(2^43 log c) / 2^43 and logctail = log c - logc rounded to double.
* log: z in [0x1.6p-1, 0x1.6p0) split likewise; invc is a double within 2^29
  ulps of 1/center such that logc = log(1/invc) rounded to double is a multiple of
  2^-43, with a small rounding error in logc.  These 128 pairs were found by that
  search and confirmed, entry by entry, to reproduce the C library's log bit for bit
  (tests/test_pmath.py checks the whole function against it); this script checks
  each pair's defining properties.

All values are computed here with Python's decimal module at 60 digits.
"""
import decimal
import struct
from decimal import Decimal

decimal.getcontext().prec = 60
N = 128

LOG_TABLE = [
    ("0x1.734f0c3e0de9fp+0", "-0x1.7cc7f79e69p-2"),
    ("0x1.713786a2ce91fp+0", "-0x1.76feec20dp-2"),
    ("0x1.6f26008fab5ap+0", "-0x1.713e31351ep-2"),
    ("0x1.6d1a61f138c7dp+0", "-0x1.6b85b382878p-2"),
    ("0x1.6b1490bc5b4d1p+0", "-0x1.65d55908078p-2"),
    ("0x1.69147332f0cbap+0", "-0x1.602d07618p-2"),
    ("0x1.6719f18224223p+0", "-0x1.5a8ca86909p-2"),
    ("0x1.6524f99a51ed9p+0", "-0x1.54f4356035p-2"),
    ("0x1.63356aa8f24c4p+0", "-0x1.4f637c36b4p-2"),
    ("0x1.614b36b9ddc14p+0", "-0x1.49da7fda85p-2"),
    ("0x1.5f66452c65c4cp+0", "-0x1.445923989a8p-2"),
    ("0x1.5d867b5912c4fp+0", "-0x1.3edf439b0b8p-2"),
    ("0x1.5babccb5b90dep+0", "-0x1.396ce448f7p-2"),
    ("0x1.59d61f2d91a78p+0", "-0x1.3401e17bdap-2"),
    ("0x1.5805612465687p+0", "-0x1.2e9e2ef468p-2"),
    ("0x1.56397cee76bd3p+0", "-0x1.2941b3830ep-2"),
    ("0x1.54725e2a77f93p+0", "-0x1.23ec58cda88p-2"),
    ("0x1.52aff42064583p+0", "-0x1.1e9e129279p-2"),
    ("0x1.50f22dbb2bddfp+0", "-0x1.1956d2b48f8p-2"),
    ("0x1.4f38f4734ded7p+0", "-0x1.141679ab9f8p-2"),
    ("0x1.4d843cfde284p+0", "-0x1.0edd094ef98p-2"),
    ("0x1.4bd3ec078a3c8p+0", "-0x1.09aa518db1p-2"),
    ("0x1.4a27fc3e0258ap+0", "-0x1.047e65263b8p-2"),
    ("0x1.4880524d48434p+0", "-0x1.feb224586fp-3"),
    ("0x1.46dce1b192d0bp+0", "-0x1.f474a7517bp-3"),
    ("0x1.453d9d3391854p+0", "-0x1.ea4443d103p-3"),
    ("0x1.43a2744b4845ap+0", "-0x1.e020d44e9bp-3"),
    ("0x1.420b54115f8fbp+0", "-0x1.d60a22977fp-3"),
    ("0x1.40782da3ef4b1p+0", "-0x1.cc00104959p-3"),
    ("0x1.3ee8f5d57fe8fp+0", "-0x1.c202956891p-3"),
    ("0x1.3d5d9a00b4ce9p+0", "-0x1.b81178d811p-3"),
    ("0x1.3bd60c010c12bp+0", "-0x1.ae2c9ccd3dp-3"),
    ("0x1.3a5242b75dab8p+0", "-0x1.a45402e129p-3"),
    ("0x1.38d22cd9fd002p+0", "-0x1.9a877681dfp-3"),
    ("0x1.3755bc5847a1cp+0", "-0x1.90c6d69483p-3"),
    ("0x1.35dce49ad36e2p+0", "-0x1.87120a645cp-3"),
    ("0x1.34679984dd44p+0", "-0x1.7d68fb4143p-3"),
    ("0x1.32f5cceffcb24p+0", "-0x1.73cb83c627p-3"),
    ("0x1.3187775a10d49p+0", "-0x1.6a39a9b376p-3"),
    ("0x1.301c8373e399p+0", "-0x1.60b3154b7ap-3"),
    ("0x1.2eb4ebb95f841p+0", "-0x1.5737d76243p-3"),
    ("0x1.2d50a0219a9d1p+0", "-0x1.4dc7b8fc23p-3"),
    ("0x1.2bef9a8b7fd2ap+0", "-0x1.4462c51d2p-3"),
    ("0x1.2a91c7a0c1babp+0", "-0x1.3b08abc83p-3"),
    ("0x1.293726014b53p+0", "-0x1.31b996b49p-3"),
    ("0x1.27dfa5757a1f5p+0", "-0x1.2875490a44p-3"),
    ("0x1.268b39b1d3bbfp+0", "-0x1.1f3b9f879ap-3"),
    ("0x1.2539d838ff5bdp+0", "-0x1.160c8252cap-3"),
    ("0x1.23eb7aac9083bp+0", "-0x1.0ce7f57f72p-3"),
    ("0x1.22a012ba940b6p+0", "-0x1.03cdc49feap-3"),
    ("0x1.2157996cc4132p+0", "-0x1.f57bdbc4b8p-4"),
    ("0x1.201201dd2fc9bp+0", "-0x1.e370896404p-4"),
    ("0x1.1ecf4494d480bp+0", "-0x1.d17983ef94p-4"),
    ("0x1.1d8f5528f6569p+0", "-0x1.bf9674ed8ap-4"),
    ("0x1.1c52311577e7cp+0", "-0x1.adc79202f6p-4"),
    ("0x1.1b17c74cb26e9p+0", "-0x1.9c0c3e7288p-4"),
    ("0x1.19e010c2c1ab6p+0", "-0x1.8a646b372cp-4"),
    ("0x1.18ab07bb670bdp+0", "-0x1.78d01b3acp-4"),
    ("0x1.1778a25efbcb6p+0", "-0x1.674f14538p-4"),
    ("0x1.1648d354c31dap+0", "-0x1.55e0e6d878p-4"),
    ("0x1.151b990275fddp+0", "-0x1.4485cdea1ep-4"),
    ("0x1.13f0ea432d24cp+0", "-0x1.333d94d6aap-4"),
    ("0x1.12c8b7210f9dap+0", "-0x1.22079f8c56p-4"),
    ("0x1.11a3028ecb531p+0", "-0x1.10e4698622p-4"),
    ("0x1.107fbda8434afp+0", "-0x1.ffa6c6ad2p-5"),
    ("0x1.0f5ee0f4e6bb3p+0", "-0x1.dda8d4a774p-5"),
    ("0x1.0e4065d2a9fcep+0", "-0x1.bbcece485p-5"),
    ("0x1.0d244632ca521p+0", "-0x1.9a1894012cp-5"),
    ("0x1.0c0a77ce2981ap+0", "-0x1.788583302cp-5"),
    ("0x1.0af2f83c636d1p+0", "-0x1.5715e67d68p-5"),
    ("0x1.09ddb98a01339p+0", "-0x1.35c8a49658p-5"),
    ("0x1.08cabaf52e7dfp+0", "-0x1.149e364154p-5"),
    ("0x1.07b9f2f4e28fbp+0", "-0x1.e72c082eb8p-6"),
    ("0x1.06ab58c358f19p+0", "-0x1.a55f152528p-6"),
    ("0x1.059eea5ecf92cp+0", "-0x1.63d62cf818p-6"),
    ("0x1.04949cdd12c9p+0", "-0x1.228fb8caap-6"),
    ("0x1.038c6c6f0ada9p+0", "-0x1.c317b20f9p-7"),
    ("0x1.02865137932a9p+0", "-0x1.419355daap-7"),
    ("0x1.0182427ea7348p+0", "-0x1.81203c2ecp-8"),
    ("0x1.008040614b195p+0", "-0x1.004097924p-9"),
    ("0x1.fe01ff726fa1ap-1", "0x1.feff3849p-9"),
    ("0x1.fa11cc261ea74p-1", "0x1.7dc41353dp-7"),
    ("0x1.f6310b081992ep-1", "0x1.3cea3c4c28p-6"),
    ("0x1.f25f63ceeadcdp-1", "0x1.b9fc11489p-6"),
    ("0x1.ee9c8039113e7p-1", "0x1.1b0d8ce11p-5"),
    ("0x1.eae8078cbb1abp-1", "0x1.58a5bd001cp-5"),
    ("0x1.e741aa29d0c9bp-1", "0x1.95c8340d88p-5"),
    ("0x1.e3a91830a99b5p-1", "0x1.d276aef578p-5"),
    ("0x1.e01e009609a56p-1", "0x1.07598e598cp-4"),
    ("0x1.dca01e577bb98p-1", "0x1.253f5e30d2p-4"),
    ("0x1.d92f20b7c9103p-1", "0x1.42edd8b38p-4"),
    ("0x1.d5cac66fb5ccep-1", "0x1.606598757cp-4"),
    ("0x1.d272caa5ede9dp-1", "0x1.7da76356ap-4"),
    ("0x1.cf26e3e6b2ccdp-1", "0x1.9ab434e1c6p-4"),
    ("0x1.cbe6da2a77902p-1", "0x1.b78c7bb0d6p-4"),
    ("0x1.c8b266d37086dp-1", "0x1.d431332e72p-4"),
    ("0x1.c5894bd5d5804p-1", "0x1.f0a3171de6p-4"),
    ("0x1.c26b533bb9f8cp-1", "0x1.067152b914p-3"),
    ("0x1.bf583eeece73fp-1", "0x1.147858292bp-3"),
    ("0x1.bc4fd75db96c1p-1", "0x1.2266ecdca3p-3"),
    ("0x1.b951e0c864a28p-1", "0x1.303d7a6c55p-3"),
    ("0x1.b65e2c5ef3e2cp-1", "0x1.3dfc33c331p-3"),
    ("0x1.b374867c9888bp-1", "0x1.4ba366b7a8p-3"),
    ("0x1.b094b211d304ap-1", "0x1.5933928d1fp-3"),
    ("0x1.adbe885f2ef7ep-1", "0x1.66acd2418fp-3"),
    ("0x1.aaf1d31603da2p-1", "0x1.740f8ec669p-3"),
    ("0x1.a82e63fd358a7p-1", "0x1.815c0f51afp-3"),
    ("0x1.a5740ef09738bp-1", "0x1.8e92954f68p-3"),
    ("0x1.a2c2a90ab4b27p-1", "0x1.9bb3602f84p-3"),
    ("0x1.a01a01393f2d1p-1", "0x1.a8bed1c2cp-3"),
    ("0x1.9d79f24db3c1bp-1", "0x1.b5b515c01dp-3"),
    ("0x1.9ae2505c7b19p-1", "0x1.c2967ccbccp-3"),
    ("0x1.9852ef297ce2fp-1", "0x1.cf635d5486p-3"),
    ("0x1.95cbaeea44b75p-1", "0x1.dc1bd3446cp-3"),
    ("0x1.934c69de74838p-1", "0x1.e8c01b8cfep-3"),
    ("0x1.90d4f2f6752e6p-1", "0x1.f5509c0179p-3"),
    ("0x1.8e6528effd79dp-1", "0x1.00e6c121fb8p-2"),
    ("0x1.8bfce9fcc007cp-1", "0x1.071b80e93dp-2"),
    ("0x1.899c0dabec30ep-1", "0x1.0d46b9e867p-2"),
    ("0x1.87427aa2317fbp-1", "0x1.13687334bdp-2"),
    ("0x1.84f00acb39a08p-1", "0x1.1980d672348p-2"),
    ("0x1.82a49e8653e55p-1", "0x1.1f8ffe0cc8p-2"),
    ("0x1.8060195f4026p-1", "0x1.2595fd76368p-2"),
    ("0x1.7e22563e0a329p-1", "0x1.2b9300914a8p-2"),
    ("0x1.7beb377dcb5adp-1", "0x1.3187210436p-2"),
    ("0x1.79baa679725c2p-1", "0x1.377266dec18p-2"),
    ("0x1.77907f2170657p-1", "0x1.3d54ffbaf3p-2"),
    ("0x1.756cadbd6130cp-1", "0x1.432eee32fep-2"),
]


def asd(u):
    return struct.unpack("<d", struct.pack("<Q", u))[0]


def asu(d):
    return struct.unpack("<Q", struct.pack("<d", d))[0]


def exp_table():
    out = []
    for k in range(N):
        exact = Decimal(2) ** (Decimal(k) / N)
        h = float(exact)
        t = float((exact - Decimal(h)) / Decimal(h))
        out.append(asu(t))
        out.append((asu(h) - (k << 45)) & 0xFFFFFFFFFFFFFFFF)
    return out


def pow_log_table():
    off = 0x3FE6955500000000
    rows = []
    for i in range(N):
        lo, hi = asd(off + (i << 45)), asd(off + ((i + 1) << 45))
        center = (lo + hi) / 2
        base = N if center < 1 else 2 * N
        j = round(base / center)
        lnc = (Decimal(base) / Decimal(j)).ln()
        logc = float(Decimal(round(lnc * 2 ** 43)) / Decimal(2) ** 43)
        rows.append((j / base, logc, float(lnc - Decimal(logc))))
    return rows


def log_table():
    off = 0x3FE6000000000000
    rows = []
    for i, (a, b) in enumerate(LOG_TABLE):
        invc, logc = float.fromhex(a), float.fromhex(b)
        lnc = -Decimal(invc).ln()
        assert float(lnc) == logc, i                              # logc = log(1/invc) rounded
        assert (768.0 + logc) - 768.0 == logc, i                  # a multiple of 2^-43
        assert abs(lnc - Decimal(logc)) < Decimal(2) ** -66, i    # small rounding error
        lo, hi = asd(off + (i << 45)), asd(off + ((i + 1) << 45))
        assert max(abs(lo * invc - 1), abs(hi * invc - 1)) < 1.0 / N, i  # c inside the subinterval
        rows.append((invc, logc))
    return rows


def main():
    print("// pmath_table.h -- generated by gen_pmath_table.py.")
    print("#pragma once")
    print("// exp: bits of T[k], bits of H[k] - (k << 45), k = 0..127")
    print("#define PMATH_EXP_TABLE \\")
    t = exp_table()
    print(",\\\n".join("  0x%016xull, 0x%016xull" % (t[2 * k], t[2 * k + 1]) for k in range(N)))
    print("// pow's logarithm: invc, logc, logctail")
    print("#define PMATH_POW_LOG_TABLE \\")
    print(",\\\n".join("  {%s, %s, %s}" % (a.hex(), b.hex(), c.hex()) for a, b, c in pow_log_table()))
    print("// log: invc, logc")
    print("#define PMATH_LOG_TABLE \\")
    print(",\\\n".join("  {%s, %s}" % (a.hex(), b.hex()) for a, b in log_table()))


if __name__ == "__main__":
    main()
