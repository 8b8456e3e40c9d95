"""Writes dg_rpc_sample.xml: a DigitalGlobe-layout .XML (tab-indented tags,
one per line, coefficient lists on the line after their <...List> tag, as
SubFunctions.cpp:1302-1582 reads them) holding made-up values. The
expected parse is listed by hand in tests/test_io.py."""
import os

coef = {
    'LINENUMCOEF': [f"{(-1)**i * (i+1) * 1.25e-3:+.6E}" for i in range(20)],
    'LINEDENCOEF': ["+1.000000E+00"] + [f"{(-1)**i * (i+2) * 2.5e-5:+.6E}" for i in range(19)],
    'SAMPNUMCOEF': [f"{(i+3) * -3.5e-4:+.6E}" for i in range(20)],
    'SAMPDENCOEF': ["+1.000000E+00"] + [f"{(i+1) * 4.0e-6:+.6E}" for i in range(19)],
}
lines = ['<?xml version="1.0" encoding="UTF-8"?>', '<isd>', '<IMD>', '\t<BAND_P>',
         '\t\t<ABSCALFACTOR>9.295654e-03</ABSCALFACTOR>', '\t\t<EFFECTIVEBANDWIDTH>2.846000e-01</EFFECTIVEBANDWIDTH>',
         '\t\t<TDILEVEL>16</TDILEVEL>', '\t</BAND_P>', '\t<IMAGE>',
         '\t\t<MEANCOLLECTEDROWGSD>0.512</MEANCOLLECTEDROWGSD>', '\t\t<MEANCOLLECTEDCOLGSD>0.498</MEANCOLLECTEDCOLGSD>',
         '\t\t<MEANCOLLECTEDGSD>0.505</MEANCOLLECTEDGSD>',
         '\t\t<MEANPRODUCTROWGSD>0.525</MEANPRODUCTROWGSD>', '\t\t<MEANPRODUCTCOLGSD>0.51</MEANPRODUCTCOLGSD>',
         '\t\t<MEANPRODUCTGSD>0.5175</MEANPRODUCTGSD>', '\t</IMAGE>', '</IMD>',
         '<RPB>', '\t<SATID>XX01</SATID>', '\t<SPECID>RPC00B</SPECID>', '\t<IMAGE>',
         '\t\t<ERRBIAS>3.75</ERRBIAS>', '\t\t<ERRRAND>0.25</ERRRAND>',
         '\t\t<LINEOFFSET>10240</LINEOFFSET>', '\t\t<SAMPOFFSET>9216</SAMPOFFSET>',
         '\t\t<LATOFFSET>-71.125</LATOFFSET>', '\t\t<LONGOFFSET>163.5</LONGOFFSET>',
         '\t\t<HEIGHTOFFSET>512</HEIGHTOFFSET>',
         '\t\t<LINESCALE>10241</LINESCALE>', '\t\t<SAMPSCALE>9217</SAMPSCALE>',
         '\t\t<LATSCALE>0.0625</LATSCALE>', '\t\t<LONGSCALE>0.1875</LONGSCALE>', '\t\t<HEIGHTSCALE>640</HEIGHTSCALE>']
for k, v in coef.items():
    lines += [f'\t\t<{k}List>', f'\t\t\t<{k}>' + ' '.join(v) + f'</{k}>', f'\t\t</{k}List>']
lines += ['\t</IMAGE>', '</RPB>', '</isd>']

if __name__ == "__main__":
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "dg_rpc_sample.xml")
    open(out, "w").write("\n".join(lines) + "\n")
