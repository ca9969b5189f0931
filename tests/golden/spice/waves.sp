Waveform sources
Vin in 0 PULSE(0 1.8 1n 0.1n 0.1n 5n 10n)
Vdc2 a 0 DC 1.2 PULSE(0 1 2n 1n 1n 3n 0)
Vsin s 0 SIN(0.5 1 1MEG 0 0 90)
Vac b 0 AC 1 0
Vbare c 0 3.3
Ipul 0 in PULSE(0 1m 0 1n 1n 2n 8n)
Isin 0 s SIN(1u 2u 10MEG)
Ishort 0 a
R1 in a 1k
R2 a s 1k
R3 s b 1k
R4 b c 1k
R5 c 0 1k
.tran 0.1n 50n 1n
.end
