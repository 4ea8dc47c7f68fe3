! ISO_C_BINDING interface to libspexhip (include/spexhip.h) for the SPEX
! Fortran host.  The shims keep the reference argument lists of the three
! call sites that change:
!   susceptibility.f:1159-1251  pair loop (HILBERT NONE)   -> spexhip_pairsum
!   susceptibility.f:778,790    wavefproducts2_mt/_pw      -> spexhip_wfprod2_mt/_pw
!   correlation.f:612-747       dielectric matrix, Minverse -> spexhip_dielec_inverse
!   correlation.f:868-872       screened interaction        -> spexhip_screen
!   correlation.f:684-690       RPA correlation integrand   -> spexhip_rpa_int
! A non-zero status stops the run with the library message, like the
! reference's Error(...) -> spex_error (util.f:480-505).
module m_spexhip
  use, intrinsic :: iso_c_binding
  implicit none
  private

  integer(c_int), parameter, public :: SPEXHIP_PREC_FP64 = 0, SPEXHIP_PREC_SPLIT_FP32 = 1, &
                                        SPEXHIP_PREC_OZAKI_INT8 = 2

  type, bind(C), public :: spexhip_mt_type
    integer(c_int) :: lcut
    type(c_ptr)    :: nindx
    integer(c_int) :: lcutm
    type(c_ptr)    :: nindxm
    integer(c_int) :: neq
    type(c_ptr)    :: integral
    integer(c_int) :: d1, d2, d3
  end type

  type, bind(C), public :: spexhip_pw_desc
    integer(c_int) :: nfft(3)
    integer(c_int) :: imati(9)
    real(c_double) :: vol
    integer(c_int) :: nkpt, maxgpt
    type(c_ptr)    :: gpt, ngpt
    integer(c_int) :: ngptm
    type(c_ptr)    :: gptm
    type(c_ptr)    :: step_r
    type(c_ptr)    :: step_c
  end type

  public :: spexhip_create, spexhip_destroy, spexhip_mt_plan_create, spexhip_pw_plan_create
  public :: spexhip_plan_destroy
  public :: spexhip_pairsum, spexhip_wfprod2_mt, spexhip_wfprod2_pw, spexhip_dielec_inverse
  public :: spexhip_minverse, spexhip_screen, spexhip_rpa_int, spexhip_pairsum_hilbert
  public :: spexhip_headwings, spexhip_dielec_inverse_gamma, spexhip_tetrahedron5
  public :: spexhip_dielecc_inverse, spexhip_tetrahedron5_equiv, spexhip_tetrahedron5_bands
  public :: spexhip_fft_setup, spexhip_step_on_grid, spexhip_alignbd, spexhip_gauss4

  interface
    integer(c_int) function spexhip_create(device, budget, ctx) bind(C, name='spexhip_create')
      import :: c_int, c_size_t, c_ptr
      integer(c_int), value :: device
      integer(c_size_t), value :: budget
      type(c_ptr) :: ctx
    end function
    integer(c_int) function spexhip_destroy(ctx) bind(C, name='spexhip_destroy')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function c_last_error(ctx, buf, len) bind(C, name='spexhip_last_error')
      import :: c_int, c_ptr, c_char, c_size_t
      type(c_ptr), value :: ctx
      character(kind=c_char) :: buf(*)
      integer(c_size_t), value :: len
    end function
    integer(c_int) function c_susceptibility_none(ctx, n, nb, cprod, ldc, nw, wght, prec, sus) &
        bind(C, name='spexhip_susceptibility_none')
      import :: c_int, c_long, c_ptr
      type(c_ptr), value :: ctx, cprod, wght, sus
      integer(c_long), value :: n, ldc
      integer(c_int), value :: nb, nw, prec
    end function
    integer(c_int) function c_susceptibility_hilbert(ctx, n, nb, cprod, ldc, nfrqs, wtetra, ldwt, &
        wghtthr, nfrqh, nfrqc, wghtSr, wghtSi, prec, sush, susc) bind(C, name='spexhip_susceptibility_hilbert')
      import :: c_int, c_long, c_ptr, c_double
      type(c_ptr), value :: ctx, cprod, wtetra, wghtSr, wghtSi, sush, susc
      integer(c_long), value :: n, ldc, ldwt
      integer(c_int), value :: nb, nfrqs, nfrqh, nfrqc, prec
      real(c_double), value :: wghtthr
    end function
    integer(c_int) function c_susceptibility_hilbert_multdiff(ctx, n, nb, cprod, ldc, nfrqs, &
        wtetra, ldwt, wghtthr, nfrqh, nfrqc, wghtSr, wghtSi, prec, frqs, nk, nocc, nunocc, &
        e_occ, e_unocc, sush, susc) bind(C, name='spexhip_susceptibility_hilbert_multdiff')
      import :: c_int, c_long, c_ptr, c_double
      type(c_ptr), value :: ctx, cprod, wtetra, wghtSr, wghtSi, frqs, e_occ, e_unocc, sush, susc
      integer(c_long), value :: n, ldc, ldwt
      integer(c_int), value :: nb, nfrqs, nfrqh, nfrqc, prec, nk, nocc, nunocc
      real(c_double), value :: wghtthr
    end function
    integer(c_int) function c_susceptibility_headwings(ctx, n, nb, cprod, ldc, dcprod, lddc, &
        nfrqh, wh, whh, nfrqc, wcr, wci, wchr, wchi, headh, wingh, headc, wingc) &
        bind(C, name='spexhip_susceptibility_headwings')
      import :: c_int, c_long, c_ptr
      type(c_ptr), value :: ctx, cprod, dcprod, wh, whh, wcr, wci, wchr, wchi, headh, wingh, &
                            headc, wingc
      integer(c_long), value :: n, ldc, lddc
      integer(c_int), value :: nb, nfrqh, nfrqc
    end function
    integer(c_int) function c_dielectric_inverse_general(ctx, nb, nw, chi, vkind, vhalf, einv) &
        bind(C, name='spexhip_dielectric_inverse_general')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, chi, vhalf, einv
      integer(c_int), value :: nb, nw, vkind
    end function
    integer(c_int) function c_tetrahedron_weights(ctx, nkpt3, rlat, nocc, nunocc, eocc, eunocc, &
        efermi, nfrqs, frqs, kpt1p, nkpt1, degocc, degunocc, wtetra) &
        bind(C, name='spexhip_tetrahedron_weights')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx, nkpt3, rlat, eocc, eunocc, frqs, kpt1p, degocc, degunocc, wtetra
      integer(c_int), value :: nocc, nunocc, nfrqs, nkpt1
      real(c_double), value :: efermi
    end function
    integer(c_int) function c_tetrahedron_weights_tetraf(ctx, nkpt3, rlat, nocc, nunocc, eocc, &
        eunocc, efermi, nfrqs, frqs, kpt1, nkpt1, kpt1p, degocc, degunocc, wtetra) &
        bind(C, name='spexhip_tetrahedron_weights_tetraf')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx, nkpt3, rlat, eocc, eunocc, frqs, kpt1, kpt1p, degocc, degunocc, &
                            wtetra
      integer(c_int), value :: nocc, nunocc, nfrqs, nkpt1
      real(c_double), value :: efermi
    end function
    integer(c_int) function c_gauss_weights(ctx, nk, nocc, nunocc, e_occ, e_unocc, efermi, &
        sigma1, sigma2, nfrqs, frqs, kpt1p, nkpt1, wtetra) bind(C, name='spexhip_gauss_weights')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx, e_occ, e_unocc, frqs, kpt1p, wtetra
      integer(c_int), value :: nk, nocc, nunocc, nfrqs, nkpt1
      real(c_double), value :: efermi, sigma1, sigma2
    end function
    integer(c_int) function c_set_alignbd(ctx, on) bind(C, name='spexhip_set_alignbd')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: on
    end function
    integer(c_int) function c_tetrahedron_weights_bands(ctx, nkpt3, rlat, nb1, ene1, deg1, b1low, &
        b1up, nb2, ene2, deg2, b2low, b2up, efermi, nfrqs, frqs, kpt1, nkpt1, kpt1p, tetraf, &
        wtetra) bind(C, name='spexhip_tetrahedron_weights_bands')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx, nkpt3, rlat, ene1, deg1, ene2, deg2, frqs, kpt1, kpt1p, wtetra
      integer(c_int), value :: nb1, b1low, b1up, nb2, b2low, b2up, nfrqs, nkpt1, tetraf
      real(c_double), value :: efermi
    end function
    integer(c_int) function c_tetra_average_equivalent(ctx, nkpt1, nocc, nunocc, nfrqs, kpt1, &
        nk, kpt1p, nsym, kptsym, ikpt, wtetra) bind(C, name='spexhip_tetra_average_equivalent')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, kpt1, kpt1p, kptsym, wtetra
      integer(c_int), value :: nkpt1, nocc, nunocc, nfrqs, nk, nsym, ikpt
    end function
    integer(c_int) function c_dielectric_inverse_gamma(ctx, nb, nw, chi, sqcoul, head, wing, &
        einv) bind(C, name='spexhip_dielectric_inverse_gamma')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, chi, sqcoul, head, wing, einv
      integer(c_int), value :: nb, nw
    end function
    integer(c_int) function spexhip_mt_plan_create(ctx, ntype, types, cent, maxlmindx, plan) &
        bind(C, name='spexhip_mt_plan_create')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, types, cent
      integer(c_int), value :: ntype, maxlmindx
      type(c_ptr) :: plan
    end function
    integer(c_int) function spexhip_pw_plan_create(ctx, desc, plan) bind(C, name='spexhip_pw_plan_create')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, desc
      type(c_ptr) :: plan
    end function
    integer(c_int) function spexhip_plan_destroy(ctx, plan) bind(C, name='spexhip_plan_destroy')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, plan
    end function
    integer(c_int) function c_wfprod_mt(ctx, plan, cmt, ncent, nband, nkpt, q, nk, k1, k2, b1lo, &
        nocc, b2lo, nunocc, cprod) bind(C, name='spexhip_wavefproducts2_mt')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, plan, cmt, q, k1, k2, cprod
      integer(c_int), value :: ncent, nband, nkpt, nk, b1lo, nocc, b2lo, nunocc
    end function
    integer(c_int) function c_wfprod_pw(ctx, plan, cpw, nband, nk, k1, k2, g2, b1lo, nocc, b2lo, &
        nunocc, cprod) bind(C, name='spexhip_wavefproducts2_pw')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, plan, cpw, k1, k2, g2, cprod
      integer(c_int), value :: nband, nk, b1lo, nocc, b2lo, nunocc
    end function
    integer(c_int) function c_dielectric_inverse(ctx, nb, nw, chi, vkind, vhalf, einv) &
        bind(C, name='spexhip_dielectric_inverse')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, chi, vhalf, einv
      integer(c_int), value :: nb, nw, vkind
    end function
    integer(c_int) function c_screened_w_packed(ctx, nb, nw, einv, vkind, vhalf, w) &
        bind(C, name='spexhip_screened_w_packed')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, einv, vhalf, w
      integer(c_int), value :: nb, nw, vkind
    end function
    integer(c_int) function c_rpa_integrand_packed(ctx, nb, nw, eps, weight, out) &
        bind(C, name='spexhip_rpa_integrand_packed')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx, eps, out
      integer(c_int), value :: nb, nw
      real(c_double), value :: weight
    end function
    integer(c_int) function c_inverse_hpd_packed(ctx, nb, count, packed) &
        bind(C, name='spexhip_inverse_hpd_packed')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, packed
      integer(c_int), value :: nb, count
    end function
    integer(c_int) function c_determine_imat(rlat, imat, imati) bind(C, name='spexhip_determine_imat')
      import :: c_int, c_ptr
      type(c_ptr), value :: rlat, imat, imati
    end function
    integer(c_int) function c_fft_grid(rlat, imat, gcutf, nfft, npoints) bind(C, name='spexhip_fft_grid')
      import :: c_int, c_ptr, c_double, c_long
      type(c_ptr), value :: rlat, imat, nfft
      real(c_double), value :: gcutf
      integer(c_long) :: npoints
    end function
    integer(c_int) function c_step_grid(nfft, imat, rlat, gmax, cstep, cext, theta) &
        bind(C, name='spexhip_step_grid')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: nfft, imat, rlat, cstep, cext, theta
      real(c_double), value :: gmax
    end function
  end interface

contains

  ! Error(...) behaviour: print the library message and stop (util.f:480-505).
  subroutine check(ctx, rc, where)
    type(c_ptr), intent(in) :: ctx
    integer(c_int), intent(in) :: rc
    character(*), intent(in) :: where
    character(kind=c_char) :: buf(512)
    character(512) :: msg
    integer :: i, irc
    if (rc == 0) return
    buf = c_null_char
    irc = c_last_error(ctx, buf, 512_c_size_t)
    msg = ' '
    do i = 1, 512
      if (buf(i) == c_null_char) exit
      msg(i:i) = buf(i)
    end do
    write(0, '(A,I0,A)') 'SPEX-ERROR (spexhip, '//where//', status ', rc, '): '//trim(msg)
    error stop 1
  end subroutine

  ! An argument shape the C side cannot check (the sizes of host arrays).
  subroutine size_error(where, msg)
    character(*), intent(in) :: where, msg
    write(0, '(A)') 'SPEX-ERROR (spexhip, '//where//'): '//msg
    error stop 1
  end subroutine

  ! Pair loop of one memory packet, HILBERT NONE (susceptibility.f:1159-1251):
  ! susceph(ij,ifrq) += sum_k wghtSr(k,ifrq) * cprod(k,i) * conjg(cprod(k,j)).
  subroutine spexhip_pairsum(ctx, n, dim, cprod, nfrqh, wghtSr, susceph, precision)
    type(c_ptr), intent(in) :: ctx
    integer, intent(in) :: n, dim, nfrqh
    complex(c_double_complex), intent(in), target :: cprod(n, dim)
    real(c_double), intent(in), target :: wghtSr(n, nfrqh)
    complex(c_double_complex), intent(inout), target :: susceph(dim*(dim+1)/2, nfrqh)
    integer(c_int), intent(in), optional :: precision
    integer(c_int) :: prec
    prec = SPEXHIP_PREC_FP64
    if (present(precision)) prec = precision
    call check(ctx, c_susceptibility_none(ctx, int(n, c_long), int(dim, c_int), c_loc(cprod), &
               int(n, c_long), int(nfrqh, c_int), c_loc(wghtSr), prec, c_loc(susceph)), 'pairsum')
  end subroutine

  ! HILBERT branch of the pair loop (susceptibility.f:1066-1251): wtetra(n,nfrqs)
  ! as filled by tetrahedron5_init (before indexing), wghtSr/wghtSi(nfrqs,
  ! nfrqh+nfrqc) from getwghtS; susceph and suscepc are accumulated.  With
  ! frqs, ene_occ(nocc,nk) (at k) and ene_unocc(nunocc,nk) (at k+q) present,
  ! MULTDIFF is ON (multdiff_ = 1, :912-925, 1193): cprod is then the packet
  ! before the scaling of its first column, n = nk*nocc*nunocc.
  subroutine spexhip_pairsum_hilbert(ctx, n, dim, cprod, nfrqs, wtetra, wghtthr, nfrqh, nfrqc, &
                                     wghtSr, wghtSi, susceph, suscepc, precision, frqs, &
                                     ene_occ, ene_unocc)
    type(c_ptr), intent(in) :: ctx
    integer, intent(in) :: n, dim, nfrqs, nfrqh, nfrqc
    complex(c_double_complex), intent(in), target :: cprod(n, dim)
    real(c_double), intent(in), target :: wtetra(n, nfrqs), wghtSr(nfrqs, nfrqh+nfrqc), &
                                          wghtSi(nfrqs, nfrqh+nfrqc)
    real(c_double), intent(in) :: wghtthr
    complex(c_double_complex), intent(inout), target :: susceph(dim*(dim+1)/2, max(1,nfrqh)), &
                                                        suscepc(dim, dim, max(1,nfrqc))
    integer(c_int), intent(in), optional :: precision
    real(c_double), intent(in), optional, target :: frqs(nfrqs)
    real(c_double), intent(in), optional, target, contiguous :: ene_occ(:,:), ene_unocc(:,:)
    integer(c_int) :: prec
    prec = SPEXHIP_PREC_FP64
    if (present(precision)) prec = precision
    if (present(frqs) .and. present(ene_occ) .and. present(ene_unocc)) then
      call check(ctx, c_susceptibility_hilbert_multdiff(ctx, int(n, c_long), int(dim, c_int), &
                 c_loc(cprod), int(n, c_long), int(nfrqs, c_int), c_loc(wtetra), int(n, c_long), &
                 wghtthr, int(nfrqh, c_int), int(nfrqc, c_int), c_loc(wghtSr), c_loc(wghtSi), &
                 prec, c_loc(frqs), int(size(ene_occ, 2), c_int), int(size(ene_occ, 1), c_int), &
                 int(size(ene_unocc, 1), c_int), c_loc(ene_occ), c_loc(ene_unocc), &
                 c_loc(susceph), c_loc(suscepc)), 'pairsum_hilbert_multdiff')
      return
    end if
    call check(ctx, c_susceptibility_hilbert(ctx, int(n, c_long), int(dim, c_int), c_loc(cprod), &
               int(n, c_long), int(nfrqs, c_int), c_loc(wtetra), int(n, c_long), wghtthr, &
               int(nfrqh, c_int), int(nfrqc, c_int), c_loc(wghtSr), c_loc(wghtSi), prec, &
               c_loc(susceph), c_loc(suscepc)), 'pairsum_hilbert')
  end subroutine

  ! Gamma-point head and wings of one memory packet (susceptibility.f:1092-1157):
  ! dcprod(n,3) after the momentum scaling (:858-873), weights per transition
  ! (wh of the wings, whh of the head, n x nfrqh; wcr/wci and wchr/wchi for the
  ! real frequencies); headh(6,nfrqh), wingh(3,dim,nfrqh), headc(3,3,nfrqc) and
  ! wingc(3,dim,2,nfrqc) are accumulated.
  subroutine spexhip_headwings(ctx, n, dim, cprod, dcprod, nfrqh, wh, whh, nfrqc, wcr, wci, &
                               wchr, wchi, headh, wingh, headc, wingc)
    type(c_ptr), intent(in) :: ctx
    integer, intent(in) :: n, dim, nfrqh, nfrqc
    complex(c_double_complex), intent(in), target :: cprod(n, dim), dcprod(n, 3)
    real(c_double), intent(in), target :: wh(n, max(1,nfrqh)), whh(n, max(1,nfrqh)), &
                                          wcr(n, max(1,nfrqc)), wci(n, max(1,nfrqc)), &
                                          wchr(n, max(1,nfrqc)), wchi(n, max(1,nfrqc))
    complex(c_double_complex), intent(inout), target :: headh(6, max(1,nfrqh)), &
        wingh(3, dim, max(1,nfrqh)), headc(3, 3, max(1,nfrqc)), wingc(3, dim, 2, max(1,nfrqc))
    call check(ctx, c_susceptibility_headwings(ctx, int(n, c_long), int(dim, c_int), c_loc(cprod), &
               int(n, c_long), c_loc(dcprod), int(n, c_long), int(nfrqh, c_int), c_loc(wh), &
               c_loc(whh), int(nfrqc, c_int), c_loc(wcr), c_loc(wci), c_loc(wchr), c_loc(wchi), &
               c_loc(headh), c_loc(wingh), c_loc(headc), c_loc(wingc)), 'headwings')
  end subroutine

  ! correlation.f:570-759 at the Gamma point (imaginary axis): suscep (packed,
  ! after the spin factor) -> dielec = eps^-1 (packed); head(6,nfreq) and
  ! wing(3,nbas,nfreq) replaced by those of the inverse; sqcoul = sqrt(coul).
  subroutine spexhip_dielec_inverse_gamma(ctx, nbas, nfreq, suscep, sqcoul, head, wing, dielec)
    type(c_ptr), intent(in) :: ctx
    integer, intent(in) :: nbas, nfreq
    complex(c_double_complex), intent(in), target :: suscep(nbas*(nbas+1)/2, nfreq)
    real(c_double), intent(in), target :: sqcoul(nbas)
    complex(c_double_complex), intent(inout), target :: head(6, nfreq), wing(3, nbas, nfreq)
    complex(c_double_complex), intent(out), target :: dielec(nbas*(nbas+1)/2, nfreq)
    call check(ctx, c_dielectric_inverse_gamma(ctx, int(nbas, c_int), int(nfreq, c_int), &
               c_loc(suscep), c_loc(sqcoul), c_loc(head), c_loc(wing), c_loc(dielec)), &
               'dielec_inverse_gamma')
  end subroutine

  ! Real frequencies (correlation.f:632-636, :701-747): dielecc = Minverse of
  ! 1 - v^1/2 suscepc v^1/2 per frequency, v^1/2 = sqcoul.
  subroutine spexhip_dielecc_inverse(ctx, nbas, nfreqc, suscepc, sqcoul, dielecc)
    type(c_ptr), intent(in) :: ctx
    integer, intent(in) :: nbas, nfreqc
    complex(c_double_complex), intent(in), target :: suscepc(nbas, nbas, nfreqc)
    real(c_double), intent(in), target :: sqcoul(nbas)
    complex(c_double_complex), intent(out), target :: dielecc(nbas, nbas, nfreqc)
    call check(ctx, c_dielectric_inverse_general(ctx, int(nbas, c_int), int(nfreqc, c_int), &
               c_loc(suscepc), 0_c_int, c_loc(sqcoul), c_loc(dielecc)), 'dielecc_inverse')
  end subroutine

  ! tetrahedron5_init (tetrahedron.f:903-1332) for kflag = 1: ene_occ(nocc,nk) at k and
  ! ene_unocc(nunocc,nk) at k+q on the full mesh nkpt3 (k = (i1*N2+i2)*N3+i3, 0-based
  ! coordinates), frqs(nfrqs) the Hilbert mesh; optional kpt1p(nk) (1-based parents,
  ! 0 = none) with nkpt1, and deg_occ(nocc,nkpt1), deg_unocc(nunocc,nkpt1) (1-based band
  ! numbers within the windows, SPEX's deg convention).  wtetra: nkr*nocc*nunocc x nfrqs
  ! (nkr = nkpt1 with kpt1p, else the mesh size), transition fastest.  kpt1(nkpt1)
  ! (1-based mesh indices) with kpt1p gives the rows of the TETRAF keyword.
  subroutine spexhip_tetrahedron5(ctx, nkpt3, rlat, nocc, nunocc, ene_occ, ene_unocc, efermi, &
                                  nfrqs, frqs, wtetra, kpt1p, nkpt1, deg_occ, deg_unocc, kpt1)
    type(c_ptr), intent(in) :: ctx
    integer, intent(in) :: nkpt3(3), nocc, nunocc, nfrqs
    real(c_double), intent(in), target :: rlat(3,3), frqs(nfrqs), &
                                          ene_occ(nocc, nkpt3(1)*nkpt3(2)*nkpt3(3)), &
                                          ene_unocc(nunocc, nkpt3(1)*nkpt3(2)*nkpt3(3))
    real(c_double), intent(in) :: efermi
    real(c_double), intent(out), target :: wtetra(*)
    integer, intent(in), optional :: kpt1p(:), nkpt1, deg_occ(:,:), deg_unocc(:,:), kpt1(:)
    integer(c_int), target :: k3(3)
    integer(c_int), allocatable, target :: kp(:), dgo(:,:), dgu(:,:), k1(:)
    type(c_ptr) :: pk, pdo, pdu
    integer(c_int) :: nk1
    k3 = nkpt3
    pk = c_null_ptr ; pdo = c_null_ptr ; pdu = c_null_ptr ; nk1 = 0
    if (present(kpt1p)) then
      ! the C side reads one parent per mesh point
      if (size(kpt1p) /= product(nkpt3)) call size_error('tetrahedron5', &
          'kpt1p must have one entry per k point of the mesh (product(nkpt3))')
      allocate(kp(size(kpt1p))) ; kp = kpt1p - 1 ; pk = c_loc(kp) ; nk1 = nkpt1
    end if
    if (present(deg_occ)) then
      allocate(dgo(size(deg_occ,1),size(deg_occ,2)), dgu(size(deg_unocc,1),size(deg_unocc,2)))
      dgo = deg_occ - 1 ; dgu = deg_unocc - 1 ; pdo = c_loc(dgo) ; pdu = c_loc(dgu)
    end if
    if (present(kpt1)) then
      if (.not. present(kpt1p)) call size_error('tetrahedron5', 'kpt1 needs kpt1p and nkpt1')
      if (size(kpt1) /= nkpt1) call size_error('tetrahedron5', 'kpt1 must have nkpt1 entries')
      allocate(k1(size(kpt1))) ; k1 = kpt1 - 1
      call check(ctx, c_tetrahedron_weights_tetraf(ctx, c_loc(k3), c_loc(rlat), int(nocc, c_int), &
                 int(nunocc, c_int), c_loc(ene_occ), c_loc(ene_unocc), efermi, &
                 int(nfrqs, c_int), c_loc(frqs), c_loc(k1), nk1, pk, pdo, pdu, c_loc(wtetra)), &
                 'tetrahedron5')
      return
    end if
    call check(ctx, c_tetrahedron_weights(ctx, c_loc(k3), c_loc(rlat), int(nocc, c_int), &
               int(nunocc, c_int), c_loc(ene_occ), c_loc(ene_unocc), efermi, int(nfrqs, c_int), &
               c_loc(frqs), pk, nk1, pdo, pdu, c_loc(wtetra)), 'tetrahedron5')
  end subroutine

  ! gauss4_init (gauss.f:120-186, kflag 1) for the GAUSS key (susceptibility.f:941-942):
  ! ene_occ(nocc,nk) at k and ene_unocc(nunocc,nk) at k+q over the full mesh, gauss the
  ! key's two widths, wtetra(nocc*nunocc*nrow, nfrqs); optional kpt1p(nk) (1-based,
  ! 0 = no parent) and nkpt1 for the rows of a reduced set.
  subroutine spexhip_gauss4(ctx, nk, nocc, nunocc, ene_occ, ene_unocc, efermi, gauss, nfrqs, &
                            frqs, wtetra, kpt1p, nkpt1)
    type(c_ptr), intent(in) :: ctx
    integer, intent(in) :: nk, nocc, nunocc, nfrqs
    real(c_double), intent(in), target :: ene_occ(nocc, nk), ene_unocc(nunocc, nk), frqs(nfrqs)
    real(c_double), intent(in) :: efermi, gauss(2)
    real(c_double), intent(out), target :: wtetra(*)
    integer, intent(in), optional :: kpt1p(:), nkpt1
    integer(c_int), allocatable, target :: kp(:)
    type(c_ptr) :: pk
    integer(c_int) :: nk1
    pk = c_null_ptr ; nk1 = 0
    if (present(kpt1p)) then
      if (size(kpt1p) /= nk) call size_error('gauss4', 'kpt1p must have one entry per k point')
      if (.not. present(nkpt1)) call size_error('gauss4', 'kpt1p needs nkpt1')
      allocate(kp(nk)) ; kp = kpt1p - 1 ; pk = c_loc(kp) ; nk1 = nkpt1
    end if
    call check(ctx, c_gauss_weights(ctx, int(nk, c_int), int(nocc, c_int), int(nunocc, c_int), &
               c_loc(ene_occ), c_loc(ene_unocc), efermi, gauss(1), gauss(2), int(nfrqs, c_int), &
               c_loc(frqs), pk, nk1, c_loc(wtetra)), 'gauss4')
  end subroutine

  ! The ALIGNBD key (getinput.f:1486): the tetrahedron shims stop before the
  ! averaging over degenerate states and equivalent k points (tetrahedron.f:1231).
  subroutine spexhip_alignbd(ctx, alignbd)
    type(c_ptr), intent(in) :: ctx
    logical, intent(in) :: alignbd
    integer(c_int) :: on
    on = 0
    if (alignbd) on = 1
    call check(ctx, c_set_alignbd(ctx, on), 'alignbd')
  end subroutine

  ! tetrahedron5_init with band windows b1low:b1up, b2low:b2up that may cut a degenerate
  ! group (tetrahedron.f:984-1012, 1191-1194, 1237-1273): ene1(nb1,nk) at k and
  ! ene2(nb2,nk) at k+q over all bands, deg1/deg2 SPEX's deg tables of the same bands
  ! (1-based, as global.f holds them), the windows 1-based; optional kpt1(nkpt1) and
  ! kpt1p(nk) (1-based, 0 = no parent) for a reduced set, tetraf for the TETRAF rows.
  ! wtetra: nkr*nocc*nunocc x nfrqs, transition fastest.
  subroutine spexhip_tetrahedron5_bands(ctx, nkpt3, rlat, ene1, deg1, b1low, b1up, ene2, deg2, &
                                        b2low, b2up, efermi, nfrqs, frqs, wtetra, kpt1, kpt1p, &
                                        tetraf)
    type(c_ptr), intent(in) :: ctx
    integer, intent(in) :: nkpt3(3), b1low, b1up, b2low, b2up, nfrqs
    real(c_double), intent(in), target :: rlat(3,3), frqs(nfrqs), ene1(:,:), ene2(:,:)
    integer, intent(in) :: deg1(:,:), deg2(:,:)
    real(c_double), intent(in) :: efermi
    real(c_double), intent(out), target :: wtetra(*)
    integer, intent(in), optional :: kpt1(:), kpt1p(:)
    logical, intent(in), optional :: tetraf
    integer(c_int), target :: k3(3)
    integer(c_int), allocatable, target :: d1(:,:), d2(:,:), k1(:), kp(:)
    real(c_double), allocatable, target :: e1(:,:), e2(:,:)
    type(c_ptr) :: pk1, pkp
    integer(c_int) :: nk1, tf
    k3 = nkpt3
    if (size(ene1,2) /= product(nkpt3) .or. size(ene2,2) /= product(nkpt3) .or. &
        any(shape(deg1) /= shape(ene1)) .or. any(shape(deg2) /= shape(ene2))) &
      call size_error('tetrahedron5_bands', 'ene/deg must be (nband, product(nkpt3)) with equal shapes')
    allocate(e1, source=ene1) ; allocate(e2, source=ene2)
    allocate(d1(size(deg1,1),size(deg1,2)), d2(size(deg2,1),size(deg2,2)))
    d1 = deg1 - 1 ; d2 = deg2 - 1
    pk1 = c_null_ptr ; pkp = c_null_ptr ; nk1 = 0 ; tf = 0
    if (present(kpt1p)) then
      if (.not. present(kpt1)) call size_error('tetrahedron5_bands', 'kpt1p needs kpt1')
      if (size(kpt1p) /= product(nkpt3)) call size_error('tetrahedron5_bands', &
          'kpt1p must have one entry per k point of the mesh (product(nkpt3))')
      allocate(k1(size(kpt1)), kp(size(kpt1p))) ; k1 = kpt1 - 1 ; kp = kpt1p - 1
      pk1 = c_loc(k1) ; pkp = c_loc(kp) ; nk1 = size(kpt1)
    end if
    if (present(tetraf)) then
      if (tetraf) tf = 1
    end if
    call check(ctx, c_tetrahedron_weights_bands(ctx, c_loc(k3), c_loc(rlat), &
               int(size(ene1,1), c_int), c_loc(e1), c_loc(d1), int(b1low-1, c_int), &
               int(b1up-1, c_int), int(size(ene2,1), c_int), c_loc(e2), c_loc(d2), &
               int(b2low-1, c_int), int(b2up-1, c_int), efermi, int(nfrqs, c_int), c_loc(frqs), &
               pk1, nk1, pkp, tf, c_loc(wtetra)), 'tetrahedron5_bands')
  end subroutine

  ! The last step of tetrahedron5_init (tetrahedron.f:1282-1318) on the
  ! wtetra of spexhip_tetrahedron5: the mean over k points of kpt1 that are
  ! equivalent under the symmetries keeping q.  1-based kpt1(nkpt1),
  ! kpt1p(nkpt) (0: no parent), kptsym(nkpt2, nsym) and ikpt, as in the
  ! reference.  SPEX allocates kptsym with nkpt2 = 2*nkpt rows when a kptadd
  ! vector is present (the shifted set follows the mesh); the averaging reads
  ! only rows of mesh points (kptsym(kpt1(k),:), kptsym(ikpt,:)), so the
  ! first nkpt rows are passed.
  subroutine spexhip_tetrahedron5_equiv(ctx, nocc, nunocc, nfrqs, kpt1, kpt1p, kptsym, ikpt, &
                                        wtetra)
    type(c_ptr), intent(in) :: ctx
    integer, intent(in) :: nocc, nunocc, nfrqs, kpt1(:), kpt1p(:), kptsym(:,:), ikpt
    real(c_double), intent(inout), target :: wtetra(*)
    integer(c_int), allocatable, target :: k1(:), kp(:), ks(:,:)
    integer :: nk
    nk = size(kpt1p)
    if (size(kptsym,1) < nk) call size_error('tetrahedron5 equivalent k', &
        'kptsym needs a row for every k point of kpt1p')
    allocate(k1(size(kpt1)), kp(nk), ks(nk, size(kptsym,2)))
    k1 = kpt1 - 1 ; kp = kpt1p - 1 ; ks = kptsym(1:nk, :) - 1
    call check(ctx, c_tetra_average_equivalent(ctx, int(size(kpt1), c_int), int(nocc, c_int), &
               int(nunocc, c_int), int(nfrqs, c_int), c_loc(k1), int(nk, c_int), &
               c_loc(kp), int(size(kptsym,2), c_int), c_loc(ks), int(ikpt-1, c_int), &
               c_loc(wtetra)), 'tetrahedron5 equivalent k')
  end subroutine

  ! wavefproducts2_mt(cprod,k1,nk,ik0,ispin1,ispin2,b1low,b1up,b2low,b2up)
  ! with 1-based k and band indices as in the reference; q = kpt(:,ik0).
  subroutine spexhip_wfprod2_mt(ctx, plan, cprod, cmt, q, k1, k2, nk, b1low, b1up, b2low, b2up)
    type(c_ptr), intent(in) :: ctx, plan
    integer, intent(in) :: nk, k1(nk), k2(nk), b1low, b1up, b2low, b2up
    complex(c_double_complex), intent(in), target :: cmt(:,:,:,:)
    real(c_double), intent(in), target :: q(3)
    complex(c_double_complex), intent(out), target :: cprod(b2low:b2up, b1low:b1up, nk, *)
    integer(c_int), target :: k1c(nk), k2c(nk)
    k1c = k1 - 1
    k2c = k2 - 1
    call check(ctx, c_wfprod_mt(ctx, plan, c_loc(cmt), int(size(cmt,2), c_int), int(size(cmt,3), c_int), &
               int(size(cmt,4), c_int), c_loc(q), int(nk, c_int), c_loc(k1c), c_loc(k2c), &
               int(b1low-1, c_int), int(b1up-b1low+1, c_int), int(b2low-1, c_int), &
               int(b2up-b2low+1, c_int), c_loc(cprod)), 'wavefproducts2_mt')
  end subroutine

  ! wavefproducts2_pw FFT path; gpt2(3,nk) is the umklapp of :547-552.
  subroutine spexhip_wfprod2_pw(ctx, plan, cprod, cpw, k1, k2, gpt2, nk, b1low, b1up, b2low, b2up)
    type(c_ptr), intent(in) :: ctx, plan
    integer, intent(in) :: nk, k1(nk), k2(nk), b1low, b1up, b2low, b2up
    integer, intent(in) :: gpt2(3, nk)
    complex(c_double_complex), intent(in), target :: cpw(:,:,:)
    complex(c_double_complex), intent(out), target :: cprod(b2low:b2up, b1low:b1up, nk, *)
    integer(c_int), target :: k1c(nk), k2c(nk), g2c(3, nk)
    k1c = k1 - 1
    k2c = k2 - 1
    g2c = gpt2
    call check(ctx, c_wfprod_pw(ctx, plan, c_loc(cpw), int(size(cpw,2), c_int), int(nk, c_int), &
               c_loc(k1c), c_loc(k2c), c_loc(g2c), int(b1low-1, c_int), int(b1up-b1low+1, c_int), &
               int(b2low-1, c_int), int(b2up-b2low+1, c_int), c_loc(cprod)), 'wavefproducts2_pw')
  end subroutine

  ! correlation.f:612-747: dielec = 1 - v^1/2 suscep v^1/2, then Minverse, per
  ! frequency.  sqcoul = sqrt(coul) (Coulomb eigenbasis).
  subroutine spexhip_dielec_inverse(ctx, nbas, nfreq, suscep, sqcoul, dielec)
    type(c_ptr), intent(in) :: ctx
    integer, intent(in) :: nbas, nfreq
    complex(c_double_complex), intent(in), target :: suscep(nbas*(nbas+1)/2, nfreq)
    real(c_double), intent(in), target :: sqcoul(nbas)
    complex(c_double_complex), intent(out), target :: dielec(nbas*(nbas+1)/2, nfreq)
    call check(ctx, c_dielectric_inverse(ctx, int(nbas, c_int), int(nfreq, c_int), c_loc(suscep), &
               0_c_int, c_loc(sqcoul), c_loc(dielec)), 'dielectric inverse')
  end subroutine

  ! correlation.f:868-872: screen = v^1/2 dielec v^1/2 per frequency, with
  ! dielec already inverted; sqcoul = sqrt(coul) (Coulomb eigenbasis).
  subroutine spexhip_screen(ctx, nbas, nfreq, dielec, sqcoul, screen)
    type(c_ptr), intent(in) :: ctx
    integer, intent(in) :: nbas, nfreq
    complex(c_double_complex), intent(in), target :: dielec(nbas*(nbas+1)/2, nfreq)
    real(c_double), intent(in), target :: sqcoul(nbas)
    complex(c_double_complex), intent(out), target :: screen(nbas*(nbas+1)/2, nfreq)
    call check(ctx, c_screened_w_packed(ctx, int(nbas, c_int), int(nfreq, c_int), c_loc(dielec), &
               0_c_int, c_loc(sqcoul), c_loc(screen)), 'screened interaction')
  end subroutine

  ! correlation.f:684-690: rpa_int(i) += sum(log(eig) + 1 - eig) of dielec(:,i)
  ! times weight = count(kptp==ikpt)/nkpt.
  subroutine spexhip_rpa_int(ctx, nbas, nfreq, dielec, weight, rpa_int)
    type(c_ptr), intent(in) :: ctx
    integer, intent(in) :: nbas, nfreq
    complex(c_double_complex), intent(in), target :: dielec(nbas*(nbas+1)/2, nfreq)
    real(c_double), intent(in) :: weight
    real(c_double), intent(inout) :: rpa_int(nfreq)
    real(c_double), target :: add(nfreq)
    call check(ctx, c_rpa_integrand_packed(ctx, int(nbas, c_int), int(nfreq, c_int), c_loc(dielec), &
               weight, c_loc(add)), 'rpa integrand')
    rpa_int = rpa_int + add
  end subroutine

  ! Minverse(dielec(:,i)) on packed Hermitian positive definite matrices.
  subroutine spexhip_minverse(ctx, mat)
    type(c_ptr), intent(in) :: ctx
    complex(c_double_complex), intent(inout), target :: mat(:)
    integer :: n
    n = nint(sqrt(0.25d0 + 2*size(mat)) - 0.5d0)
    call check(ctx, c_inverse_hpd_packed(ctx, int(n, c_int), 1_c_int, c_loc(mat)), 'Minverse')
  end subroutine

  ! getinput.f:2472-2473 (call determine_imat; call fft_init(gcutf)): the
  ! basis change imat/imati of fft.f:163-238 and the grid nfft of fft.f:36-51
  ! and 101-158 for the reciprocal basis rlat(3,3).  The C side keeps
  ! imat/imati row-major, hence the transposes.
  subroutine spexhip_fft_setup(rlat, gcutf, nfft, imat, imati)
    real(c_double), intent(in), target :: rlat(3, 3)
    real(c_double), intent(in) :: gcutf
    integer, intent(out) :: nfft(3), imat(3, 3), imati(3, 3)
    integer(c_int), target :: im(3, 3), imi(3, 3), nf(3)
    integer(c_long) :: npts
    integer(c_int) :: rc
    rc = c_determine_imat(c_loc(rlat), c_loc(im), c_loc(imi))
    if (rc /= 0) call size_error('fft_setup', 'determine_imat failed (degenerate lattice)')
    imat = transpose(im)   ! im(j,i) holds imat(i,j)
    imati = transpose(imi)
    rc = c_fft_grid(c_loc(rlat), c_loc(im), gcutf, c_loc(nf), npts)
    if (rc /= 0) call size_error('fft_setup', 'more than 2^16 points in one direction')
    nfft = nf
  end subroutine

  ! The step function on the FFT grid as wavefproducts.f:576-590 builds it for
  ! FULLPW: theta(0:nfft(1)-1,0:nfft(2)-1,0:nfft(3)-1) from cstep(-c1:c1,
  ! -c2:c2,-c3:c3) cut at gmax = 2*gcut+gcutm.  Pass theta as the plan's
  ! step_c.
  subroutine spexhip_step_on_grid(nfft, imat, rlat, gmax, cstep, theta)
    integer, intent(in) :: nfft(3), imat(3, 3)
    real(c_double), intent(in), target :: rlat(3, 3)
    real(c_double), intent(in) :: gmax
    complex(c_double_complex), intent(in), target :: cstep(:, :, :)
    complex(c_double_complex), intent(out), target :: theta(0:nfft(1)-1, 0:nfft(2)-1, 0:nfft(3)-1)
    integer(c_int), target :: im(3, 3), nf(3), cext(3)
    integer :: i
    do i = 1, 3
      if (mod(size(cstep, i), 2) /= 1) call size_error('step_on_grid', 'cstep extents must be odd')
      cext(i) = (size(cstep, i) - 1) / 2
    end do
    nf = nfft
    im = transpose(imat)
    if (c_step_grid(c_loc(nf), c_loc(im), c_loc(rlat), gmax, c_loc(cstep), c_loc(cext), &
                    c_loc(theta)) /= 0) &
      call size_error('step_on_grid', 'a frequency inside gmax lies outside cstep')
  end subroutine

end module m_spexhip
