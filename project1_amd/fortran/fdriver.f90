! Fortran host exercising the ISO_C_BINDING boundary: one packet of the
! HILBERT NONE pair sum and the dielectric inverse, compared with a plain
! Fortran pair loop written the way susceptibility.f:1183-1240 does it; the
! screened interaction, the RPA integrand, the HILBERT pair sum (also with
! MULTDIFF), the Gamma-point head and wings, the equivalent-k averaging and
! the GAUSS weights likewise.
! Exit status 0 on agreement (normwise relative error < 1e-12).
program fdriver
  use, intrinsic :: iso_c_binding
  use m_spexhip
  implicit none
  integer, parameter :: n = 96, nb = 40, nw = 3, npack = nb*(nb+1)/2
  complex(c_double_complex), target :: cprod(n, nb), sus(npack, nw), ref(npack, nw), einv(npack, nw)
  real(c_double), target :: wght(n, nw), sq(nb)
  real(c_double) :: re(n), im(n), err, nrm, x
  type(c_ptr) :: ctx
  integer :: i, j, k, ij, f, rc

  do j = 1, nb
    do k = 1, n
      x = 0.37d0*k + 1.13d0*j
      cprod(k, j) = cmplx(sin(x), cos(1.7d0*x), kind=c_double_complex) / sqrt(dble(nb))
    end do
  end do
  do f = 1, nw
    do k = 1, n
      wght(k, f) = -2d0*(0.5d0 + mod(k, 7)) / ((0.3d0*f)**2 + (0.5d0 + mod(k, 7))**2) / 8
    end do
  end do
  do i = 1, nb
    sq(i) = sqrt(0.1d0 + i)
  end do

  ! plain Fortran pair loop (reference order: j outer, i <= j, one dot per frequency)
  ref = 0
  do j = 1, nb
    do i = 1, j
      do k = 1, n
        re(k) = real(cprod(k, i) * conjg(cprod(k, j)))
        im(k) = aimag(cprod(k, i) * conjg(cprod(k, j)))
      end do
      ij = j*(j-1)/2 + i
      do f = 1, nw
        ref(ij, f) = ref(ij, f) + cmplx(dot_product(re, wght(:, f)), dot_product(im, wght(:, f)), &
                                        kind=c_double_complex)
      end do
    end do
  end do

  rc = spexhip_create(0_c_int, 0_c_size_t, ctx)
  if (rc /= 0) then
    write(0, *) 'spexhip_create failed', rc
    stop 2
  end if
  sus = 0
  call spexhip_pairsum(ctx, n, nb, cprod, nw, wght, sus)
  err = maxval(abs(sus - ref))
  nrm = maxval(abs(ref))
  write(*, '(A,ES10.3)') 'fortran pairsum max-rel error: ', err / nrm
  if (err / nrm > 1d-12) stop 3

  sus = 2 * sus  ! spin factor (susceptibility.f:1336)
  call spexhip_dielec_inverse(ctx, nb, nw, sus, sq, einv)
  ! eps^-1 of the first frequency times eps must be the identity
  block
    complex(c_double_complex) :: e(nb, nb), ei(nb, nb), prod(nb, nb)
    ij = 0
    do j = 1, nb
      do i = 1, j
        ij = ij + 1
        e(i, j) = -sq(i) * sq(j) * sus(ij, 1)
        e(j, i) = conjg(e(i, j))
        ei(i, j) = einv(ij, 1)
        ei(j, i) = conjg(ei(i, j))
      end do
      e(j, j) = e(j, j) + 1
    end do
    prod = matmul(ei, e)
    do i = 1, nb
      prod(i, i) = prod(i, i) - 1
    end do
    write(*, '(A,ES10.3)') 'fortran eps^-1 * eps - 1 max: ', maxval(abs(prod))
    if (maxval(abs(prod)) > 1d-10) stop 4
  end block
  ! screened interaction: screen = sqcoul einv sqcoul
  block
    complex(c_double_complex) :: scr(nb*(nb+1)/2, nw)
    real(c_double) :: d
    call spexhip_screen(ctx, nb, nw, einv, sq, scr)
    d = 0
    ij = 0
    do j = 1, nb
      do i = 1, j
        ij = ij + 1
        d = max(d, abs(scr(ij, 1) - sq(i) * einv(ij, 1) * sq(j)))
      end do
    end do
    write(*, '(A,ES10.3)') 'fortran screen max error: ', d / maxval(abs(scr(:, 1)))
    if (d / maxval(abs(scr(:, 1))) > 1d-13) stop 5
  end block
  ! RPA integrand of a diagonal dielectric matrix: sum(log d + 1 - d) / 8
  block
    complex(c_double_complex) :: dg(nb*(nb+1)/2, 1)
    real(c_double) :: rint(1), want, dd
    dg = 0
    want = 0
    do j = 1, nb
      dd = 1 + real(j, c_double) / nb
      dg(j*(j+1)/2, 1) = dd
      want = want + (log(dd) + 1 - dd) / 8
    end do
    rint = 0
    call spexhip_rpa_int(ctx, nb, 1, dg, 0.125d0, rint)
    write(*, '(A,ES10.3)') 'fortran rpa_int rel error: ', abs(rint(1) - want) / abs(want)
    if (abs(rint(1) - want) > 1d-12 * abs(want)) stop 6
  end block
  ! HILBERT branch: spectral function per pair on a mesh of nfs points, then
  ! the Hilbert sums (susceptibility.f:1793-1849), into susceph (imaginary
  ! frequencies) and suscepc (real ones, :1231-1236)
  block
    integer, parameter :: nfs = 10, nfh = 2, nfc = 2
    real(c_double), target :: wt(n, nfs), wsr(nfs, nfh+nfc), wsi(nfs, nfh+nfc)
    complex(c_double_complex), target :: sh(npack, nfh), sc(nb, nb, nfc)
    complex(c_double_complex) :: rh(npack, nfh), rcx(nb, nb, nfc)
    integer :: s
    do s = 1, nfs
      do k = 1, n
        wt(k, s) = 0
        if (mod(k + s, 4) == 0) wt(k, s) = 0.1d0 + 0.01d0 * mod(3*k + s, 17)
      end do
      do f = 1, nfh + nfc
        wsr(s, f) = -1d0 / (0.2d0*s + 0.5d0*f)
        wsi(s, f) = 0.05d0 * sin(0.7d0*s + f)
      end do
    end do
    call hilbert_literal(cprod, wt, wsr, wsi, nfh, .false., wsr(:, 1), rh, rcx)
    sh = 0
    sc = 0
    call spexhip_pairsum_hilbert(ctx, n, nb, cprod, nfs, wt, 1d-6, nfh, nfc, wsr, wsi, sh, sc)
    err = max(maxval(abs(sh - rh)) / maxval(abs(rh)), maxval(abs(sc - rcx)) / maxval(abs(rcx)))
    write(*, '(A,ES10.3)') 'fortran pairsum_hilbert max-rel error: ', err
    if (err > 1d-12) stop 7
    ! MULTDIFF ON (susceptibility.f:912-925, 1193, 1817-1824): n = 2 k points
    ! x 4 occupied x 12 unoccupied bands; the mesh starts at 0
    block
      integer, parameter :: nko = 2, no = 4, nu = 12
      real(c_double), target :: eo(no, nko), eu(nu, nko), fq(nfs)
      complex(c_double_complex) :: cm(n, nb)
      integer :: ik, b1, b2
      do s = 1, nfs
        fq(s) = 0.15d0 * (s - 1)
      end do
      do ik = 1, nko
        do b1 = 1, no
          eo(b1, ik) = -0.3d0 * b1 - 0.1d0 * ik
        end do
        do b2 = 1, nu
          eu(b2, ik) = 0.05d0 * b2 + 0.2d0 * ik
        end do
      end do
      cm = cprod
      k = 0
      do ik = 1, nko
        do b1 = 1, no
          do b2 = 1, nu
            k = k + 1
            cm(k, 1) = cm(k, 1) * (eu(b2, ik) - eo(b1, ik))
          end do
        end do
      end do
      call hilbert_literal(cm, wt, wsr, wsi, nfh, .true., fq, rh, rcx)
      sh = 0
      sc = 0
      call spexhip_pairsum_hilbert(ctx, n, nb, cprod, nfs, wt, 1d-6, nfh, nfc, wsr, wsi, sh, sc, &
                                   frqs=fq, ene_occ=eo, ene_unocc=eu)
      err = max(maxval(abs(sh - rh)) / maxval(abs(rh)), maxval(abs(sc - rcx)) / maxval(abs(rcx)))
      write(*, '(A,ES10.3)') 'fortran pairsum_hilbert MULTDIFF max-rel error: ', err
      if (err > 1d-12) stop 13
    end block
  end block

  ! Gamma point: head and wings of the packet against plain loops written the
  ! way susceptibility.f:1096-1127 does them (HILBERT NONE), then the head and
  ! wings of the inverse: its eps^-1 must equal the plain inverse of the chi
  ! whose first row and column are zero (correlation.f:597-603).
  block
    complex(c_double_complex), target :: dc(n, 3), hh(6, nw), wh(3, nb, nw), rhh(6, nw), &
        rwh(3, nb, nw), hc(3, 3, 1), wc(3, nb, 2, 1), chz(npack, nw), e1(npack, nw), e2(npack, nw)
    real(c_double), target :: dummy(n, 1)
    integer :: k1, k2
    do k2 = 1, 3
      do k = 1, n
        x = 0.21d0*k + 0.9d0*k2
        dc(k, k2) = cmplx(cos(x), sin(0.6d0*x), kind=c_double_complex)
      end do
    end do
    rhh = 0
    rwh = 0
    do k2 = 1, 3
      do k1 = 1, k2
        re = real(dc(:, k1) * conjg(dc(:, k2)))
        im = aimag(dc(:, k1) * conjg(dc(:, k2)))
        ij = (k2-1)*k2/2 + k1
        do f = 1, nw
          rhh(ij, f) = rhh(ij, f) + cmplx(dot_product(re, wght(:, f)), dot_product(im, wght(:, f)), &
                                          kind=c_double_complex)
        end do
      end do
      do i = 1, nb
        re = real(cprod(:, i) * conjg(dc(:, k2)))
        im = aimag(cprod(:, i) * conjg(dc(:, k2)))
        do f = 1, nw
          rwh(k2, i, f) = rwh(k2, i, f) + cmplx(dot_product(re, wght(:, f)), &
                                                dot_product(im, wght(:, f)), kind=c_double_complex)
        end do
      end do
    end do
    hh = 0
    wh = 0
    dummy = 0
    call spexhip_headwings(ctx, n, nb, cprod, dc, nw, wght, wght, 0, dummy, dummy, dummy, dummy, &
                           hh, wh, hc, wc)
    err = max(maxval(abs(hh - rhh)) / maxval(abs(rhh)), maxval(abs(wh - rwh)) / maxval(abs(rwh)))
    write(*, '(A,ES10.3)') 'fortran headwings max-rel error: ', err
    if (err > 1d-12) stop 8
    chz = sus
    do i = 1, nb
      chz(1 + i*(i-1)/2, :) = 0
    end do
    call spexhip_dielec_inverse(ctx, nb, nw, chz, sq, e1)
    hh = hh * 1d-2
    wh = wh * 1d-2
    call spexhip_dielec_inverse_gamma(ctx, nb, nw, sus, sq, hh, wh, e2)
    err = maxval(abs(e2 - e1)) / maxval(abs(e1))
    write(*, '(A,ES10.3)') 'fortran dielec_inverse_gamma body max-rel error: ', err
    if (err > 1d-12 .or. any(wh(:, 1, :) /= 0)) stop 9
  end block
  ! Averaging over equivalent k points (tetrahedron.f:1282-1318) through the
  ! shim with a kptsym table of nkpt2 = 2*nkpt rows, as SPEX allocates it when
  ! a kptadd vector is present; the second half maps into the shifted set
  ! (indices > nkpt).  The shim must hand over only the mesh rows, so its
  ! result equals the call with the first nkpt rows alone.  With
  ! SPEXHIP_FDRIVER_OUT set, tables and weights are written there for a
  ! comparison with the oracle.
  block
    integer, parameter :: nk = 10, nsym = 4, nkpt1 = 7, no = 3, nu = 5, nf = 6, ik = 4
    integer :: ks2(2*nk, nsym), kp1(nkpt1), kpp(nk), s, u
    real(c_double), target :: w0(nkpt1*no*nu, nf), w1(nkpt1*no*nu, nf), w2(nkpt1*no*nu, nf)
    character(len=512) :: outdir
    integer :: ln
    do s = 1, nsym
      do k = 1, nk
        ks2(k, s) = mod(7*k + 3*s + k*s, nk) + 1
        ks2(nk + k, s) = nk + mod(5*k + s, nk) + 1
      end do
    end do
    ks2(1:nk, 1) = [(k, k = 1, nk)]
    ks2(ik, :) = ik
    kp1 = [3, 1, 8, 5, 10, 6, 2]
    do k = 1, nk
      kpp(k) = mod(3*k, nkpt1 + 1)
    end do
    do f = 1, nf
      do i = 1, nkpt1*no*nu
        w0(i, f) = 0.5d0 + 0.3d0*sin(0.1d0*i + 0.7d0*f)
      end do
    end do
    w1 = w0
    call spexhip_tetrahedron5_equiv(ctx, no, nu, nf, kp1, kpp, ks2, ik, w1)
    w2 = w0
    call spexhip_tetrahedron5_equiv(ctx, no, nu, nf, kp1, kpp, ks2(1:nk, :), ik, w2)
    write(*, '(A,ES10.3)') 'fortran tetrahedron5_equiv (2*nkpt kptsym) change: ', &
        maxval(abs(w1 - w0))
    if (any(w1 /= w2) .or. maxval(abs(w1 - w0)) == 0) stop 10
    call get_environment_variable('SPEXHIP_FDRIVER_OUT', outdir, ln)
    if (ln > 0) then
      open(newunit=u, file=outdir(1:ln)//'/equiv.bin', access='stream', form='unformatted', &
           status='replace')
      write(u) nk, nsym, nkpt1, no, nu, nf, ik, kp1, kpp, ks2, w0, w1
      close(u)
    end if
  end block
  ! Tetrahedron weights with band windows that cut degenerate groups, through
  ! the shim with SPEX's 1-based deg tables; written out for the oracle.
  block
    integer, parameter :: nk = 8, nb = 4, nf = 10
    integer :: n3(3), d1(nb, nk), d2(nb, nk), b, u
    real(c_double), target :: rl(3,3), e1(nb, nk), e2(nb, nk), fr(nf), wt(nk*2*2, nf)
    character(len=512) :: outdir
    integer :: ln
    n3 = [2, 2, 2]
    rl = 0 ; rl(1,1) = 1 ; rl(2,2) = 1.1d0 ; rl(3,3) = 0.9d0
    do k = 1, nk
      do b = 1, nb
        e1(b, k) = -1.2d0 + 0.25d0*b + 0.05d0*sin(1.3d0*k + b)
        e2(b, k) = 0.3d0 + 0.2d0*b + 0.04d0*cos(0.7d0*k*b)
      end do
      e1(3, k) = e1(2, k)  ! occupied pair 2-3 on every k
      e2(2, k) = e2(1, k)  ! unoccupied pair 1-2 on every k
      d1(:, k) = [1, 3, 2, 4]
      d2(:, k) = [2, 1, 3, 4]
    end do
    do f = 1, nf
      fr(f) = 0.2d0 * (1.3d0**(f-1) - 1)
    end do
    ! windows 3:4 (cuts the occupied pair) and 2:3 (cuts the unoccupied pair)
    call spexhip_tetrahedron5_bands(ctx, n3, rl, e1, d1, 3, 4, e2, d2, 2, 3, 0d0, nf, fr, wt)
    write(*, '(A,ES10.3)') 'fortran tetrahedron5_bands weight sum: ', sum(wt)
    if (.not. sum(wt) > 0) stop 11
    call get_environment_variable('SPEXHIP_FDRIVER_OUT', outdir, ln)
    if (ln > 0) then
      open(newunit=u, file=outdir(1:ln)//'/tetra_bands.bin', access='stream', &
           form='unformatted', status='replace')
      write(u) rl, e1, e2, fr, d1, d2, wt
      close(u)
    end if
  end block
  ! GAUSS weights (gauss4_init) through the shim against the loop of
  ! gauss.f:149-185 written out here, with the compiler's erf; a reduced set
  ! of 3 rows, one mesh point without a parent.
  block
    integer, parameter :: nk = 6, no = 2, nu = 3, nf = 9
    real(c_double), target :: eo(no, nk), eu(nu, nk), fr(nf), wt(no*nu*3, nf), rw(nu, no, 3, nf)
    real(c_double) :: gs(2), a, r, r1, r2, d, df, gfac, g2fac, fac, pi
    integer :: kp(nk), b1, b2, p
    pi = 4 * atan(1d0)
    gs = [0.03d0, 0.05d0]
    do k = 1, nk
      do b1 = 1, no
        eo(b1, k) = -0.6d0 + 0.2d0*b1 + 0.03d0*sin(1.1d0*k*b1)
      end do
      do b2 = 1, nu
        eu(b2, k) = -0.02d0 + 0.3d0*b2 + 0.04d0*cos(0.9d0*k + b2)
      end do
    end do
    do f = 1, nf
      fr(f) = 0.05d0 * (1.4d0**(f-1) - 1)
    end do
    kp = [1, 2, 2, 0, 3, 1]
    call spexhip_gauss4(ctx, nk, no, nu, eo, eu, 0d0, gs, nf, fr, wt, kp, 3)
    r = 0
    do f = 2, nf - 1
      r = r + (fr(f) - fr(f-1)) / (fr(f+1) - fr(f-1))
      a = r / (f - 1)
    end do
    gfac = 1 / (sqrt(2d0) * gs(2))
    g2fac = 1 / (sqrt(2*pi) * gs(2))
    fac = sqrt(0.5d0*pi) * gs(2)
    rw = 0
    do k = 1, nk
      p = kp(k) ; if (p == 0) cycle
      do b1 = 1, no
        r1 = (1 + erf((0d0 - eo(b1, k)) / (sqrt(2d0) * gs(1)))) / 2
        do b2 = 1, nu
          r2 = (1 + erf((eu(b2, k) - 0d0) / (sqrt(2d0) * gs(1)))) / 2
          d = eu(b2, k) - eo(b1, k)
          do f = 1, nf
            if (fr(f) == 0) cycle
            if (f /= nf) then ; df = a * (fr(f+1) - fr(f))
            else              ; df = (1 - a) * (fr(f) - fr(f-1))
            end if
            rw(b2, b1, p, f) = rw(b2, b1, p, f) + r1 * r2 * g2fac * fac * &
                (erf(gfac * (fr(f) + df - d)) - erf(gfac * (fr(f) - df - d))) / (2*df)
          end do
        end do
      end do
    end do
    rw = rw / nk
    err = maxval(abs(wt - reshape(rw, [no*nu*3, nf]))) / maxval(abs(rw))
    write(*, '(A,ES10.3)') 'fortran gauss4 max-rel error: ', err
    if (err > 1d-12 .or. any(wt(:, 1) /= 0)) stop 14
  end block
  ! FFT grid setup (getinput.f:2472-2473) for a simple cubic reciprocal lattice
  ! of spacing 1.1 written in a skewed basis: determine_imat must make the
  ! basis orthogonal again, and the grid of the 8.3 sphere is then that of
  ! the cubic lattice (half-width 7 -> 15 points).  With SPEXHIP_FDRIVER_OUT
  ! set, rlat, nfft, imat and imati go there for the oracle.
  block
    real(c_double) :: rlat(3, 3), r1(3, 3), c
    integer :: nfft(3), imat(3, 3), imati(3, 3), u, ln, p, q
    character(len=512) :: outdir
    rlat = 1.1d0 * reshape([1d0, 0d0, 1d0, 1d0, 1d0, 1d0, 0d0, -1d0, 1d0], [3, 3])
    call spexhip_fft_setup(rlat, 8.3d0, nfft, imat, imati)
    r1 = matmul(rlat, dble(imat))
    c = 0
    do p = 1, 3
      do q = p + 1, 3
        c = c + dot_product(r1(:, p), r1(:, q))**2
      end do
    end do
    write(*, '(A,3I4,A,ES10.3)') 'fortran fft_setup nfft:', nfft, '  non-orthogonality: ', c
    if (any(matmul(imat, imati) /= reshape([1, 0, 0, 0, 1, 0, 0, 0, 1], [3, 3]))) stop 11
    if (c > 1d-20 .or. any(nfft /= 15)) stop 12
    call get_environment_variable('SPEXHIP_FDRIVER_OUT', outdir, ln)
    if (ln > 0) then
      open(newunit=u, file=outdir(1:ln)//'/fftgrid.bin', access='stream', form='unformatted', &
           status='replace')
      write(u) rlat, nfft, imat, imati
      close(u)
    end if
  end block
  rc = spexhip_destroy(ctx)
  write(*, '(A)') 'fortran driver OK'

contains

  ! The HILBERT pair loop written the way susceptibility.f:1183-1236 and
  ! spectral_function (:1793-1849) do it; md: MULTDIFF's division of the
  ! spectral function of a pair with l of its indices on the first basis
  ! function by fq(s)**l where fq(s) > 0 (:1193, :1817-1824).
  subroutine hilbert_literal(cprod, wt, wsr, wsi, nfh, md, fq, rh, rcx)
    complex(c_double_complex), intent(in) :: cprod(:, :)
    real(c_double), intent(in) :: wt(:, :), wsr(:, :), wsi(:, :), fq(:)
    integer, intent(in) :: nfh
    logical, intent(in) :: md
    complex(c_double_complex), intent(out) :: rh(:, :), rcx(:, :, :)
    real(c_double) :: re(size(cprod, 1)), im(size(cprod, 1)), reS(size(wt, 2)), &
                      imS(size(wt, 2)), s1, s2, c1, c2
    complex(c_double_complex) :: cdum, cdum1
    integer :: i, j, k, s, f, ij, l, n, nb, nfs, nfc
    n = size(cprod, 1)
    nb = size(cprod, 2)
    nfs = size(wt, 2)
    nfc = size(wsr, 2) - nfh
    rh = 0
    rcx = 0
    do j = 1, nb
      do i = 1, j
        do k = 1, n
          re(k) = real(cprod(k, i) * conjg(cprod(k, j)))
          im(k) = aimag(cprod(k, i) * conjg(cprod(k, j)))
        end do
        do s = 1, nfs
          reS(s) = 0
          imS(s) = 0
          do k = 1, n
            if (wt(k, s) > 1d-6) then
              reS(s) = reS(s) + re(k) * wt(k, s)
              imS(s) = imS(s) + im(k) * wt(k, s)
            end if
          end do
        end do
        if (md) then
          l = count([i, j] == 1)
          do s = 1, nfs
            if (l > 0 .and. fq(s) > 0) then
              reS(s) = reS(s) / fq(s)**l
              imS(s) = imS(s) / fq(s)**l
            end if
          end do
        end if
        ij = j*(j-1)/2 + i
        do f = 1, nfh
          rh(ij, f) = rh(ij, f) + cmplx(dot_product(reS, wsr(:, f)), dot_product(imS, wsr(:, f)), &
                                        kind=c_double_complex)
        end do
        do f = 1, nfc
          s1 = dot_product(reS, wsr(:, nfh+f))
          s2 = dot_product(imS, wsr(:, nfh+f))
          c1 = dot_product(reS, wsi(:, nfh+f))
          c2 = dot_product(imS, wsi(:, nfh+f))
          cdum = cmplx(s1, c1, kind=c_double_complex)
          cdum1 = cmplx(-c2, s2, kind=c_double_complex)
          rcx(i, j, f) = rcx(i, j, f) + cdum + cdum1
          if (i /= j) rcx(j, i, f) = rcx(j, i, f) + cdum - cdum1
        end do
      end do
    end do
  end subroutine

end program fdriver
